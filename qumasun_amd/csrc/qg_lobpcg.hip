// qg_lobpcg.hip — device-resident block LOBPCG (placeholder while the driver is written).
#include "qg_internal.hpp"

using namespace qg;

extern "C" {

void qg_lobpcg_default_opts(qg_lobpcg_opts* o) {
    if (!o) return;
    o->precision = QG_PREC_F64;
    o->max_iter = 100;
    o->tol = 0.0;
    o->refresh_period = 10;
    o->rollback_limit = 1.0e-3;
    o->precond = QG_PRECOND_NONE;
    o->precond_shift = 1.0;
    o->verbose = 0;
}

int qg_lobpcg_d_multi(qg_ham* h, const qg_lobpcg_opts* opts, int nbands, double* eig, double* X,
                      double* P, int scf_step, qg_lobpcg_stats* stats) {
    return guarded([&] { throw Error{QG_ERR_ARG, "LOBPCG driver not built in this revision"}; });
}

}
