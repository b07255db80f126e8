// SME level 1 numerics (hydrostatic reconstruction, Rusanov, path-conservative moments).
#pragma once
#include "Model.H"
template <typename T>
struct Numerics : SmeNumerics2D<T, 1> {};
