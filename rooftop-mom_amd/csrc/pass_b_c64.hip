// Pass B of the matvec in complex64, every transform plan (the kernels of pass_b.inl).
#include "pass_b.inl"

namespace rt {
template void launch_pass_b<float>(const Product &, const PlanTables &, const PlanTables &, cplx<float> *, const cplx<float> *, int, int, int, hipStream_t, const int *, int, const KernelTimer *);
}  // namespace rt
