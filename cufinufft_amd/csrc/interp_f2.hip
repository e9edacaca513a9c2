// Interp and coordinate-derivative kernels for float, 2D (one
// translation unit per precision, dimension and side, so the ns
// instantiations compile in parallel).
#include "spreadinterp_impl.inl"
CFA_INSTANTIATE_INTERP(float, 2)
