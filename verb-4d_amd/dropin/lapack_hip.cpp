// Drop-in for the LAPACK routines MatrixSolver binds
// (src/MatrixSolver.h:243-248): dgbtrf_ / dgbtrs_ with the Fortran ABI and
// ILP64 `long` integers the reference declares, computed on the GPU
// (LAPACK pivoting, dgbtf2 storage convention). Linking this object in place
// of a LAPACK library moves every band factorisation the reference's own
// Lapack() performs onto the device.
#include <cstdint>

#include "verb_dropin.h"

extern "C" void dgbtrf_(long* m, long* n, long* kl, long* ku, double* ab, long* ldab, long* ipiv, long* info) {
    int rc = verb_dgbtrf(verb_dropin::ctx(), *m, *n, *kl, *ku, ab, *ldab, reinterpret_cast<int64_t*>(ipiv));
    *info = rc;  // > 0: zero pivot column (like dgbtrf), < 0: argument/device error
}

extern "C" void dgbtrs_(char const* trans, long* n, long* kl, long* ku, long* nrhs, double* ab, long* ldab,
                        long* ipiv, double* b, long* ldb, long* info) {
    if (*trans != 'N' && *trans != 'n') {
        *info = -1;
        return;
    }
    *info = 0;
    for (long r = 0; r < *nrhs; r++) {
        int rc = verb_dgbtrs(verb_dropin::ctx(), *n, *kl, *ku, ab, *ldab, reinterpret_cast<const int64_t*>(ipiv),
                             b + r * (*ldb));
        if (rc != 0) {
            *info = rc;
            return;
        }
    }
}
