// Interp and coordinate-derivative kernels for double, 1D (one
// translation unit per precision, dimension and side, so the ns
// instantiations compile in parallel).
#include "spreadinterp_impl.inl"
CFA_INSTANTIATE_INTERP(double, 1)
