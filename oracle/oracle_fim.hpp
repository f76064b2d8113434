// TEST INFRASTRUCTURE ONLY: CPU restatement of IsoT_FIM's property pass and
// assembly (oracle + CPU baseline).  Not part of the product.
#ifndef ORACLE_FIM_HPP
#define ORACLE_FIM_HPP

#include "oracle_flash.hpp"

namespace oracle {

void CalFlash(const ocpg_eos& eos, const ocpg_bulk& bk, int nthreads);
void InitFlash(const ocpg_eos& eos, const ocpg_bulk& bk, const double* zi, int nthreads);
void CalKrPc(const ocpg_satfunc& sat, const ocpg_bulk& bk, int nthreads);
void CalRock(const ocpg_rock& rock, const ocpg_bulk& bk, int nthreads);
void CalRes(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt, double* resAbs,
            int32_t* upblock, double* upblock_Rho, double* upblock_Velocity, ocpg_resnorm* norms);
void BsrPattern(int64_t nb, const ocpg_conn& conn, int32_t* IA, int32_t* JA);
void AssembleBulks(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt,
                   const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                   const int32_t* JA, double* val);

}  // namespace oracle

#endif
