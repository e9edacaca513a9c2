// Spread kernels and plan groups for double, 1D (one translation unit
// per precision, dimension and side, so the ns instantiations compile in
// parallel).
#include "spreadinterp_impl.inl"
CFA_INSTANTIATE_SPREAD(double, 1)
