// SME level 0 on a 2-D horizontal mesh with a diffusive face term: the model surface of
// ../sme0 plus Model::diffusive_flux, the optional hook a generated model may emit and the
// solver adds after the flux of every face (TransportStep.hpp:10-18, 468-491, 533-550).
// The term is nu_h (d/dx + d/dy) of each momentum from the face-averaged cell gradients
// (the hook receives no normal; as in the reference the value times the face area goes to the
// face's cells with opposite signs).  nu_h = p[3].
#pragma once
#include "../sme2d.H"
template <typename T>
struct Model : SmeModel2D<T, 0, 0, 1, 2, 4> {
    static constexpr bool has_diffusion = true;
    static const std::vector<std::string> parameter_names() { return {"g", "nu", "lambda", "nu_h"}; }
    static const std::vector<T> default_parameters() { return {9.81, 1.0e-3, 0.1, 2.0e-3}; }
    PORTABLE_FN static inline SimpleArray<T, 4> diffusive_flux(const T* Q, const T* Qaux, const T* gradQ,
                                                               const T* p) {
        SimpleArray<T, 4> r;
        r[0] = 0;
        r[1] = 0;
        r[2] = p[3] * (gradQ[2 * 2 + 0] + gradQ[2 * 2 + 1]);
        r[3] = p[3] * (gradQ[3 * 2 + 0] + gradQ[3 * 2 + 1]);
        return r;
    }
};
