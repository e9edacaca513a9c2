// Spread/interp kernels for float, 3D (one translation unit per precision
// and dimension so the ns instantiations compile in parallel).
#include "spreadinterp_impl.inl"
CFA_INSTANTIATE_SPREADINTERP(float, 3)
