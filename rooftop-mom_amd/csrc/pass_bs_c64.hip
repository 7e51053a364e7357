// Pass B in three launches (B1 forward y transforms, B2 multiply, B3 inverse) in complex64, every transform plan
// (the kernels of pass_b_split.inl).
#include "pass_b_split.inl"

namespace rt {
template void launch_pass_b_split<float>(const Product &, const PlanTables &, const PlanTables &, cplx<float> *, cplx<float> *,
                                     const cplx<float> *, int, hipStream_t, const int *, int);
}  // namespace rt
