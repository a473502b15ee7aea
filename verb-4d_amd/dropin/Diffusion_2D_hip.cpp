// Drop-in for src/Diffusion_2D.cpp: the same signatures
// (src/Diffusion_2D.h:86-111, plain and LU_CACHING variants), the local
// (E, alpha) step computed by the gfx950 kernels through verb_diffusion_2d.
// Matrix2D / Matrix1D keep their contiguous plane_array, which is handed
// over as a plain pointer.
#include "Diffusion_2D.h"

#include "verb_dropin.h"

static bool diffusion_2d_hip(Matrix2D<double>& psd, const Matrix2D<double>& x, const Matrix2D<double>& y, int x_size,
                             int y_size, const Matrix1D<double>& x_LBC, const Matrix1D<double>& x_UBC,
                             const Matrix1D<double>& y_LBC, const Matrix1D<double>& y_UBC,
                             BoundaryConditionType x_LBC_type, BoundaryConditionType x_UBC_type,
                             BoundaryConditionType y_LBC_type, BoundaryConditionType y_UBC_type,
                             const Matrix2D<double>& Dxx, const Matrix2D<double>& Dyy, const Matrix2D<double>& Dxy,
                             const Matrix2D<double>& Dyx, const Matrix2D<double>& G, const Matrix2D<double>& Sources,
                             const Matrix2D<double>& Losses, double dt, double sub_dt, double* lu_cache,
                             long* index_cache, bool recompute_lu) {
    int rc = verb_diffusion_2d(
        verb_dropin::ctx(), psd.plane_array, x.plane_array, y.plane_array, x_size, y_size, x_LBC.plane_array,
        x_UBC.plane_array, y_LBC.plane_array, y_UBC.plane_array, (int)x_LBC_type, (int)x_UBC_type, (int)y_LBC_type,
        (int)y_UBC_type, Dxx.plane_array, Dyy.plane_array, Dxy.plane_array, Dyx.plane_array, G.plane_array,
        Sources.plane_array, Losses.plane_array, dt, sub_dt, lu_cache, reinterpret_cast<int64_t*>(index_cache),
        recompute_lu ? 1 : 0);
    if (rc > 0) {
        Logger::error << "Error computing LU decomposition\n";
        exit(EXIT_FAILURE);
    }
    verb_dropin::check(rc, "Diffusion_2D");
    return true;
}

#ifdef LU_CACHING
bool Diffusion_2D(Matrix2D<double>& psd, const Matrix2D<double>& x, const Matrix2D<double>& y, int x_size, int y_size,
                  const Matrix1D<double>& x_LBC, const Matrix1D<double>& x_UBC, const Matrix1D<double>& y_LBC,
                  const Matrix1D<double>& y_UBC, BoundaryConditionType x_LBC_type, BoundaryConditionType x_UBC_type,
                  BoundaryConditionType y_LBC_type, BoundaryConditionType y_UBC_type, const Matrix2D<double>& Dxx,
                  const Matrix2D<double>& Dyy, const Matrix2D<double>& Dxy, const Matrix2D<double>& Dyx,
                  const Matrix2D<double>& G, const Matrix2D<double>& Sources, const Matrix2D<double>& Losses,
                  double dt, double sub_dt, double* lu_cache, long* index_cache, bool recompute_lu) {
    return diffusion_2d_hip(psd, x, y, x_size, y_size, x_LBC, x_UBC, y_LBC, y_UBC, x_LBC_type, x_UBC_type,
                            y_LBC_type, y_UBC_type, Dxx, Dyy, Dxy, Dyx, G, Sources, Losses, dt, sub_dt, lu_cache,
                            index_cache, recompute_lu);
}
#else
bool Diffusion_2D(Matrix2D<double>& psd, const Matrix2D<double>& x, const Matrix2D<double>& y, int x_size, int y_size,
                  const Matrix1D<double>& x_LBC, const Matrix1D<double>& x_UBC, const Matrix1D<double>& y_LBC,
                  const Matrix1D<double>& y_UBC, BoundaryConditionType x_LBC_type, BoundaryConditionType x_UBC_type,
                  BoundaryConditionType y_LBC_type, BoundaryConditionType y_UBC_type, const Matrix2D<double>& Dxx,
                  const Matrix2D<double>& Dyy, const Matrix2D<double>& Dxy, const Matrix2D<double>& Dyx,
                  const Matrix2D<double>& G, const Matrix2D<double>& Sources, const Matrix2D<double>& Losses,
                  double dt, double sub_dt) {
    return diffusion_2d_hip(psd, x, y, x_size, y_size, x_LBC, x_UBC, y_LBC, y_UBC, x_LBC_type, x_UBC_type,
                            y_LBC_type, y_UBC_type, Dxx, Dyy, Dxy, Dyx, G, Sources, Losses, dt, sub_dt, nullptr,
                            nullptr, true);
}
#endif
