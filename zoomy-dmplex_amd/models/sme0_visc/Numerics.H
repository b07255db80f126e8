// SME level 0 on a 2-D horizontal mesh with a diffusive face term: Numerics<T> surface (the
// scheme of ../sme2d.H).
#pragma once
#include "Model.H"
template <typename T>
struct Numerics : SmeNumerics2D<T, 0, 0, 1, 2, 4> {};
