// Pass A of the matvec in complex128, every transform plan (the kernels of pass_a.inl).
#include "pass_a.inl"

namespace rt {
template void launch_pass_a<double>(const Product &, const PlanTables &, const c128 *, cplx<double> *, int, int, hipStream_t, const int *, int, const KernelTimer *);
}  // namespace rt
