// Transform-pass launchers instantiated for double.
#include "passes_impl.hpp"

namespace lg {

template bool choice_ok<double>(int, const PassChoice&, int, int);
template PassChoice default_choice<double>(int, int, int);
template PassChoice choice_for<double>(int, int, int);
template void launch_rows_r2c<double>(KernelTimer&, const FftPlan&, int, const double* const*,
                                   const double*, long, int, int, long, cplx<double>* const*, long,
                                   int, long, int, int, int, int, hipStream_t,
                                   const PassChoice*, const double*, int, int);
template void launch_rows_c2r<double>(KernelTimer&, const FftPlan&, int, const cplx<double>* const*,
                                   long, int, long, double* const*, long, int, int, long, int, int,
                                   int, int, hipStream_t, const PassChoice*, const int*,
                                   double* const*, double);
template void launch_cols<double, -1>(KernelTimer&, ColArgs<double>&, int, int, int, int, hipStream_t,
                                   const PassChoice*);
template void launch_cols<double, +1>(KernelTimer&, ColArgs<double>&, int, int, int, int, hipStream_t,
                                   const PassChoice*);

std::vector<std::pair<int, int>> pass_splits(int N) { return splits_for(N); }
template bool launch_lag_rows<double>(KernelTimer&, const LagRowArgs<double>&, int, hipStream_t);
template bool lag_rows_ok<double>(int, int);
template bool launch_rows_r2c_prod<double>(KernelTimer&, const FftPlan&, const ProdSrc<double>&, long,
                                          int, int, long, cplx<double>* const*, long, long, int,
                                          int, int, int, hipStream_t);
template bool prod_rows_ok<double>(int, int);
template bool launch_lag_prep_rows<double>(KernelTimer&, const LagPrepArgs<double>&, int, hipStream_t);

}  // namespace lg
