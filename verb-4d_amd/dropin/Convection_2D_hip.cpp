// Drop-in for src/Convection_2D.cpp (src/Convection_2D.h:98-108): the
// (V,K) plane advanced by the gfx950 convection kernel through
// verb_convection_2d. The SAVE_PSD_LOST_CONV variant (lost-PSD bookkeeping)
// has no GPU counterpart.
#include "Convection_2D.h"

#include "verb_dropin.h"

#ifdef SAVE_PSD_LOST_CONV
#error "the HIP drop-in of Convection_2D does not implement SAVE_PSD_LOST_CONV"
#endif

bool Convection_2D(Matrix2D<double>& PSD_PR, const Matrix2D<double>& P, const Matrix2D<double>& R, int P_size,
                   int R_size, const Matrix1D<double>& P_LBC, const Matrix1D<double>& P_UBC,
                   const Matrix1D<double>& R_LBC, const Matrix1D<double>& R_UBC, BoundaryConditionType P_LBC_type,
                   BoundaryConditionType P_UBC_type, BoundaryConditionType R_LBC_type,
                   BoundaryConditionType R_UBC_type, const Matrix2D<double>& VP, const Matrix2D<double>& VR,
                   const Matrix2D<double>& Sources, const Matrix2D<double>& Losses, const Matrix2D<double>& G_conv,
                   double dt_total) {
    int rc = verb_convection_2d(verb_dropin::ctx(), PSD_PR.plane_array, P.plane_array, R.plane_array, P_size, R_size,
                                P_LBC.plane_array, P_UBC.plane_array, R_LBC.plane_array, R_UBC.plane_array,
                                (int)P_LBC_type, (int)P_UBC_type, (int)R_LBC_type, (int)R_UBC_type, VP.plane_array,
                                VR.plane_array, Sources.plane_array, Losses.plane_array, G_conv.plane_array, dt_total);
    verb_dropin::check(rc, "Convection_2D");
    return true;
}
