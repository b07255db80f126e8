// VAM level-1 predictor on a 2-D horizontal mesh (the pressure-free hyperbolic part of the
// vertically-averaged-and-moment model that ChorinVAMSolver advances, ChorinVAM.hpp:475-484).
// The reference ships no predictor header for it (its create_vam_model.py needs zoomy_core),
// so this build uses: SME level 1 for [b, h, q_x_0, q_x_1, q_y_0, q_y_1] plus the vertical
// velocity moments [r_0, r_1] advected with the mean velocity, Newtonian slip friction on the
// horizontal momenta.  State and parameter layout match what ChorinOps.H reads:
//   Q = [b, h, q_x_0, q_x_1, q_y_0, q_y_1, r_0, r_1]   (ChorinOps.H:3)
//   p = [g, rho, nu, lambda]                            (p[1] = rho, ChorinOps.H:103-138)
#pragma once
#define ZOOMY_MODEL_IS_VAM 1
#include "../sme2d.H"

template <typename T>
struct Model : SmeModel2D<T, 1, 2, 2, 3, 4> {
    static const std::vector<std::string> parameter_names() { return {"g", "rho", "nu", "lambda"}; }
    static const std::vector<T> default_parameters() { return {9.81, 1.0, 1.0e-3, 0.1}; }
};
