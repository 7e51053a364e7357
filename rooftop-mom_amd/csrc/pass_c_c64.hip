// Pass C of the matvec in complex64, every transform plan (the kernels of pass_c.inl).
#include "pass_c.inl"

namespace rt {
template void launch_pass_c<float>(const Product &, const PlanTables &, const cplx<float> *, c128 *, int, int, hipStream_t, const int *, int, const KernelTimer *);
}  // namespace rt
