// Transform-pass launchers instantiated for float.
#include "passes_impl.hpp"

namespace lg {

template bool choice_ok<float>(int, const PassChoice&, int, int);
template PassChoice default_choice<float>(int, int, int);
template PassChoice choice_for<float>(int, int, int);
template void launch_rows_r2c<float>(KernelTimer&, const FftPlan&, int, const float* const*,
                                   const float*, long, int, int, long, cplx<float>* const*, long,
                                   int, long, int, int, int, int, hipStream_t,
                                   const PassChoice*, const double*, int, int);
template void launch_rows_c2r<float>(KernelTimer&, const FftPlan&, int, const cplx<float>* const*,
                                   long, int, long, float* const*, long, int, int, long, int, int,
                                   int, int, hipStream_t, const PassChoice*, const int*,
                                   float* const*, float);
template void launch_cols<float, -1>(KernelTimer&, ColArgs<float>&, int, int, int, int, hipStream_t,
                                   const PassChoice*);
template void launch_cols<float, +1>(KernelTimer&, ColArgs<float>&, int, int, int, int, hipStream_t,
                                   const PassChoice*);
template bool launch_lag_rows<float>(KernelTimer&, const LagRowArgs<float>&, int, hipStream_t);
template bool lag_rows_ok<float>(int, int);
template bool launch_rows_r2c_prod<float>(KernelTimer&, const FftPlan&, const ProdSrc<float>&, long,
                                          int, int, long, cplx<float>* const*, long, long, int,
                                          int, int, int, hipStream_t);
template bool prod_rows_ok<float>(int, int);
template bool launch_lag_prep_rows<float>(KernelTimer&, const LagPrepArgs<float>&, int, hipStream_t);

}  // namespace lg
