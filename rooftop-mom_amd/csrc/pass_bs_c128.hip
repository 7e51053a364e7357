// Pass B in three launches (B1 forward y transforms, B2 multiply, B3 inverse) in complex128, every transform plan
// (the kernels of pass_b_split.inl).
#include "pass_b_split.inl"

namespace rt {
template void launch_pass_b_split<double>(const Product &, const PlanTables &, const PlanTables &, cplx<double> *, cplx<double> *,
                                     const cplx<double> *, int, hipStream_t, const int *, int);
}  // namespace rt
