// Drop-in for src/Convection_1D_ULTIMATE_QUICKEST6.cpp
// (src/Convection_1D_ULTIMATE_QUICKEST6.h:71-79) through
// verb_convection_1d_quickest6.
#include "Convection_1D_ULTIMATE_QUICKEST6.h"

#include "verb_dropin.h"

bool Convection_1D_ULTIMATE_QUICKEST6(Matrix1D<double>& PSD, const Matrix1D<double>& x, int x_size, double x_LBC,
                                      double x_UBC, BoundaryConditionType x_LBC_type,
                                      BoundaryConditionType x_UBC_type, const Matrix1D<double>& Ux,
                                      double dt_total) {
    int rc = verb_convection_1d_quickest6(verb_dropin::ctx(), PSD.plane_array, x.plane_array, x_size, x_LBC, x_UBC,
                                          (int)x_LBC_type, (int)x_UBC_type, Ux.plane_array, dt_total);
    if (rc == VERB_ERR_COURANT) {
        Logger::error << "max(CourNum) > 1, calculation can't be performed." << std::endl;
        exit(EXIT_FAILURE);
    }
    verb_dropin::check(rc, "CONV_1D");
    return true;
}
