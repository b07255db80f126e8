// SME level 0 on a 2-D horizontal mesh: Model<T> surface (see ../sme2d.H for the equations).
#pragma once
#include "../sme2d.H"
template <typename T>
struct Model : SmeModel2D<T, 0> {};
