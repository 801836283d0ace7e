// C entry point onto the reference's own Quadrature class -- TEST
// INFRASTRUCTURE ONLY.  Compiled together with the reference sources
// Quadrature.cc, Global.cc and Assert.cc (which build on their own) into
// oracle/_ref/libref_quadrature.so by `make -C oracle ref`; the tests use it
// to pin the oracle's and the engine's quadrature tables.
#include "Quadrature.hh"

extern "C" int ref_quadrature(int sn, double *xi, double *eta, double *mu,
                              double *wt)
{
    Quadrature q(static_cast<UINT>(sn));
    const UINT na = g_nAngles;
    for (UINT a = 0; a < na; a++) {
        xi[a] = q.getXi(a);
        eta[a] = q.getEta(a);
        mu[a] = q.getMu(a);
        wt[a] = q.getWt(a);
    }
    return static_cast<int>(na);
}
