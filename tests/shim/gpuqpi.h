// Test-only stand-in for calcQPI's GPU plug-in interface (include/gpuqpi.h:18-32
// of the reference): the pure-virtual class the shim implements, with the
// entry points the reference's tmatrix calls.
#pragma once
#include <ostream>
#include <vector>

#include "gsl_standin.h"
#include "idl.h"

class GPUQPI
{
public:
    virtual void wannierldos(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0) = 0;
    virtual void wannierjosephson(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0,
                                  gsl_complex tip) = 0;
    virtual void wannierldoslist(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0) = 0;
    virtual void spf(std::vector<std::vector<gsl_matrix_complex *> > &g0) = 0;
    virtual void uspf(std::vector<std::vector<gsl_matrix_complex *> > &g0) = 0;
    virtual void retrieveResult(idl &map, size_t layer) = 0;
    virtual void retrieveResult(double *ldos) = 0;
    virtual void retrieveResult(std::vector<std::vector<double> > &ldos) = 0;
    virtual void printinfo(std::ostream &os) = 0;
    virtual ~GPUQPI() {}
};
