// VAM level-1 predictor numerics (see Model.H): SME hydrostatic-reconstruction Rusanov with
// the path-conservative moment product; r_0, r_1 are upwinded with the same wave speed.
#pragma once
#include "Model.H"
template <typename T>
struct Numerics : SmeNumerics2D<T, 1, 2, 2, 3, 4> {};
