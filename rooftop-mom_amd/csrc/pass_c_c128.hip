// Pass C of the matvec in complex128, every transform plan (the kernels of pass_c.inl).
#include "pass_c.inl"

namespace rt {
template void launch_pass_c<double>(const Product &, const PlanTables &, const cplx<double> *, c128 *, int, int, hipStream_t, const int *, int, const KernelTimer *);
}  // namespace rt
