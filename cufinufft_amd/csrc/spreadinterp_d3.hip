// Spread/interp kernels for double, 3D (one translation unit per precision
// and dimension so the ns instantiations compile in parallel).
#include "spreadinterp_impl.inl"
CFA_INSTANTIATE_SPREADINTERP(double, 3)
