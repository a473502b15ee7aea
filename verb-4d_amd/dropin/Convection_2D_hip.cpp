// Drop-in for src/Convection_2D.cpp: the (V,K) plane advanced by the gfx950
// convection kernel through verb_convection_2d_ex. Both signatures of the
// reference are provided, as its build options select them
// (src/Convection_2D.h:84-109): the SAVE_PSD_LOST_CONV one accumulates the
// PSD taken by the losses into PSD_lost_PR, and FAST_CONVECTION selects the
// smaller of the two outer sub-step counts (src/Convection_2D.cpp:80-84).
#include "Convection_2D.h"

#include "verb_dropin.h"

namespace {

#ifdef FAST_CONVECTION
constexpr int kFast = 1;
#else
constexpr int kFast = 0;
#endif

void convection(Matrix2D<double>& PSD_PR, double* lost, const Matrix2D<double>& P, const Matrix2D<double>& R,
                int P_size, int R_size, const Matrix1D<double>& P_LBC, const Matrix1D<double>& P_UBC,
                const Matrix1D<double>& R_LBC, const Matrix1D<double>& R_UBC, BoundaryConditionType P_LBC_type,
                BoundaryConditionType P_UBC_type, BoundaryConditionType R_LBC_type, BoundaryConditionType R_UBC_type,
                const Matrix2D<double>& VP, const Matrix2D<double>& VR, const Matrix2D<double>& Sources,
                const Matrix2D<double>& Losses, const Matrix2D<double>& G_conv, double dt_total) {
    const verb_convection_options opt{lost, kFast};
    int rc = verb_convection_2d_ex(verb_dropin::ctx(), PSD_PR.plane_array, P.plane_array, R.plane_array, P_size,
                                   R_size, P_LBC.plane_array, P_UBC.plane_array, R_LBC.plane_array,
                                   R_UBC.plane_array, (int)P_LBC_type, (int)P_UBC_type, (int)R_LBC_type,
                                   (int)R_UBC_type, VP.plane_array, VR.plane_array, Sources.plane_array,
                                   Losses.plane_array, G_conv.plane_array, dt_total, &opt);
    verb_dropin::check(rc, "Convection_2D");
}

}  // namespace

#ifdef SAVE_PSD_LOST_CONV
bool Convection_2D(Matrix2D<double>& PSD_PR, Matrix2D<double>& PSD_lost_PR, const Matrix2D<double>& P,
                   const Matrix2D<double>& R, int P_size, int R_size, const Matrix1D<double>& P_LBC,
                   const Matrix1D<double>& P_UBC, const Matrix1D<double>& R_LBC, const Matrix1D<double>& R_UBC,
                   BoundaryConditionType P_LBC_type, BoundaryConditionType P_UBC_type,
                   BoundaryConditionType R_LBC_type, BoundaryConditionType R_UBC_type, const Matrix2D<double>& VP,
                   const Matrix2D<double>& VR, const Matrix2D<double>& Sources, const Matrix2D<double>& Losses,
                   const Matrix2D<double>& G_conv, double dt_total) {
    convection(PSD_PR, PSD_lost_PR.plane_array, P, R, P_size, R_size, P_LBC, P_UBC, R_LBC, R_UBC, P_LBC_type,
               P_UBC_type, R_LBC_type, R_UBC_type, VP, VR, Sources, Losses, G_conv, dt_total);
    return true;
}
#else
bool Convection_2D(Matrix2D<double>& PSD_PR, const Matrix2D<double>& P, const Matrix2D<double>& R, int P_size,
                   int R_size, const Matrix1D<double>& P_LBC, const Matrix1D<double>& P_UBC,
                   const Matrix1D<double>& R_LBC, const Matrix1D<double>& R_UBC, BoundaryConditionType P_LBC_type,
                   BoundaryConditionType P_UBC_type, BoundaryConditionType R_LBC_type,
                   BoundaryConditionType R_UBC_type, const Matrix2D<double>& VP, const Matrix2D<double>& VR,
                   const Matrix2D<double>& Sources, const Matrix2D<double>& Losses, const Matrix2D<double>& G_conv,
                   double dt_total) {
    convection(PSD_PR, nullptr, P, R, P_size, R_size, P_LBC, P_UBC, R_LBC, R_UBC, P_LBC_type, P_UBC_type,
               R_LBC_type, R_UBC_type, VP, VR, Sources, Losses, G_conv, dt_total);
    return true;
}
#endif
