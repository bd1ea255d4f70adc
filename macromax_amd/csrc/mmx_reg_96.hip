// Register-resident pass kernels for axes of 96 points (see mmx_reglaunch.hpp).
#include "mmx_reglaunch.hpp"

namespace mmx {
int launch_reg_96(int kind, bool three, int chi_kind, const TileCfg& t, const PassArgs<float>& a,
                  hipStream_t stream) {
    return launch_reg_len<96>(kind, three, chi_kind, t, a, stream);
}
}  // namespace mmx
