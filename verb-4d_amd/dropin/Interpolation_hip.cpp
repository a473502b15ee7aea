// Drop-in for src/Interpolation.cpp (src/Interpolation.h:25): Cubic1D runs
// on the GPU through verb_cubic1d. As in the reference, `values` is replaced
// by the interpolant at new_grid and the returned Matrix1D is empty; the
// time loop's `PSD_L = Cubic1D(old, PSD_L, new)` (src/VERB4D_Solver.cpp:597)
// then reads PSD_L's array, which keeps the new values. Input the
// reference's interpolator rejects with an exception is rejected with one
// here too.
#include <stdexcept>

#include "Interpolation.h"
#include "verb_dropin.h"

Matrix1D<double> Cubic1D(Matrix1D<double>& old_grid, Matrix1D<double>& values, Matrix1D<double>& new_grid) {
    if (new_grid.size_q1 != old_grid.size_q1) throw("Only interpolation to the same grid size is supported.");
    const int rc = verb_cubic1d(verb_dropin::ctx(), (int)old_grid.size_q1, old_grid.plane_array, values.plane_array,
                                new_grid.plane_array);
    if (rc == VERB_ERR_DOMAIN) throw std::runtime_error(verb_ctx_last_error(verb_dropin::ctx()));
    verb_dropin::check(rc, "Cubic1D");
    return Matrix1D<double>();
}
