// SME level 0 on a 2-D horizontal mesh: Numerics<T> surface (see ../sme2d.H for the scheme).
#pragma once
#include "Model.H"
template <typename T>
struct Numerics : SmeNumerics2D<T, 0> {};
