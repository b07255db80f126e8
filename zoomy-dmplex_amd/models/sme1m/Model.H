// SME level 1 with Manning friction and a per-cell Manning coefficient (aux slot 1).
#pragma once
#include "../sme_manning.H"
template <typename T>
struct Model : SmeManning2D<T, 1> {};
