// Pass B of the matvec in complex128, every transform plan (the kernels of pass_b.inl).
#include "pass_b.inl"

namespace rt {
template void launch_pass_b<double>(const Product &, const PlanTables &, const PlanTables &, cplx<double> *, const cplx<double> *, int, int, int, hipStream_t, const int *, int, const KernelTimer *);
}  // namespace rt
