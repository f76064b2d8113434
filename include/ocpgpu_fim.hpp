/* C++ host shim over the C-ABI (ocpgpu.h) with the method set of the
 * reference's IsoT_FIM (include/OCPFluidMethod.hpp, src/OCPFluidMethod.cpp:
 * 836-2857) and the sequence IsothermalSolver drives it through
 * (src/IsothermalSolver.cpp:79-155):
 *
 *   InitReservoir  -> IsoT_FIM::InitReservoir  (InitRock, InitFlash, CalKrPc)
 *   Prepare        -> IsoT_FIM::Prepare        (per-step well values)
 *   AssembleMat    -> IsoT_FIM::AssembleMat    (bulk + well blocks, rhs)
 *   UpdateProperty -> IsoT_FIM::UpdateProperty (CalFlash, CalKrPc, CalRock,
 *                                               well CalTrans/CalFlux, CalRes)
 *   UpdatePropertyAssemble -> both of the above, overlapped on the device
 *   GetSolution    -> IsoT_FIM::GetSolution    (Newton update of the state)
 *   NRChange       -> the dSNR / NRdSmax / maxNRdSSP that FinishNR reads
 *   UpdateLastTimeStep / ResetToLastTimeStep
 *
 * The caller keeps its Bulk / BulkConn / Well arrays in the reference's own
 * cell-major layout and hands pointers to them (ocpg_bulk, ocpg_conn,
 * ocpg_wells).  The cell state stays on the GPU between calls; Upload and
 * Download move it.  Errors throw ocpg::Error carrying the C-ABI status (the
 * reference aborts instead).  One host thread per object.
 */
#ifndef OCPGPU_FIM_HPP
#define OCPGPU_FIM_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ocpgpu.h"

namespace ocpg {

class Error : public std::runtime_error {
public:
    Error(const char* what, int code) : std::runtime_error(std::string(what) + " failed with status " +
                                                           std::to_string(code)),
                                        code(code) {}
    int code;
};

inline void check(int rc, const char* what)
{
    if (rc != OCPG_OK) throw Error(what, rc);
}

/* The assembled linear system in LinearSystem's / FASP dBSRmat's layout. */
struct BsrSystem {
    int64_t rows = 0, nnzb = 0;
    int blockDim = 0;
    std::vector<int32_t> IA, JA;
    std::vector<double> val, rhs;
};

class IsoT_FIM_GPU {
public:
    /* MixtureComp's deck data (ParamReservoir::comsParam), the saturation
     * tables and the rock model; device = GPU ordinal.  method
     * OCPG_METHOD_FIMN gives IsoT_FIMn's method set (src/OCPFluidMethod.cpp:
     * 2962-4017) behind the same members: the bulk table then also binds
     * nj, res_n and resPc. */
    IsoT_FIM_GPU(const ocpg_eos_deck& deck, const ocpg_satfunc& sat, const ocpg_rock& rock, int device = 0,
                 int method = OCPG_METHOD_FIM)
    {
        check(ocpg_eos_from_deck(&deck, &eos_), "ocpg_eos_from_deck");
        check(ocpg_create(&ctx_, &eos_, &sat, &rock, device), "ocpg_create");
        if (method != OCPG_METHOD_FIM) {
            const int rc = ocpg_set_method(ctx_, method);
            if (rc != OCPG_OK) {
                ocpg_destroy(ctx_);
                ctx_ = nullptr;
                check(rc, "ocpg_set_method");
            }
        }
    }
    IsoT_FIM_GPU(const IsoT_FIM_GPU&) = delete;
    IsoT_FIM_GPU& operator=(const IsoT_FIM_GPU&) = delete;
    ~IsoT_FIM_GPU() { ocpg_destroy(ctx_); }

    /* IsoT_FIM::Setup / AllocateReservoir: cells and connections (BulkConn
     * order: bId < eId, sorted by (bId, eId)). */
    void Setup(int64_t nb, const ocpg_conn& conn)
    {
        check(ocpg_set_cells(ctx_, nb), "ocpg_set_cells");
        check(ocpg_set_conn(ctx_, &conn), "ocpg_set_conn");
    }

    /* IsoT_FIM::InitReservoir: InitRock (CalRock), InitFlash (InitFlashFIM
     * from the feed zi, nb x NC row-major; InitFlashFIMn for IsoT_FIMn,
     * src/OCPFluidMethod.cpp:2975-2989) and CalKrPc; the result becomes
     * the last time step.  bk supplies P, T, S (water column), v, depth and
     * poroInit and receives the flash outputs. */
    void InitReservoir(ocpg_bulk& bk, const double* zi)
    {
        Upload(bk);
        check(ocpg_calrock(ctx_), "ocpg_calrock");
        check(ocpg_init_flash_fim(ctx_, zi), "ocpg_init_flash_fim");
        check(ocpg_calkrpc(ctx_), "ocpg_calkrpc");
        check(ocpg_update_last_step(ctx_), "ocpg_update_last_step");
        Download(bk);
    }

    /* IsoT_FIM::Prepare: the wells of the coming time step (structure and
     * the per-step values AllWells::PrepareWell computes). */
    void Prepare(const ocpg_wells& wells)
    {
        check(ocpg_set_wells(ctx_, &wells), "ocpg_set_wells");
    }
    /* New BHPs (IsoT_FIM::GetSolution's well update) or controls. */
    void UpdateWellState(const ocpg_wells& wells) { check(ocpg_set_well_state(ctx_, &wells), "ocpg_set_well_state"); }
    /* The wells' setup data (Well::Setup: depth, injected phase and
     * temperature, initial mode), once after Prepare(wells); from then on
     * AllWells::PrepareWell can run on the device. */
    void SetWellSetup(const ocpg_well_setup& su) { check(ocpg_set_well_setup(ctx_, &su), "ocpg_set_well_setup"); }
    /* AllWells::PrepareWell on the device (CalTrans, CaldG, the injectors'
     * molar density, CheckOptMode); IsoT_FIM::Prepare then calls CalRes. */
    void PrepareWells() { check(ocpg_prepare_wells(ctx_), "ocpg_prepare_wells"); }
    /* AllWells::CheckP at the end of a converged Newton iteration
     * (IsoT_FIM::FinishNR): OCPG_WELL_SUCCESS, _NEGATIVE_PRESSURE or _SWITCH. */
    int CheckWells()
    {
        int32_t st = 0;
        check(ocpg_check_wells(ctx_, &st), "ocpg_check_wells");
        return st;
    }
    /* Bulk::CheckNi and Bulk::CheckP (IsoT_FIM::UpdateProperty): 0 or
     * OCPG_BULK_NEGATIVE_COMPONENTS_MOLES / OCPG_BULK_NEGATIVE_PRESSURE. */
    int CheckBulk()
    {
        int32_t st = 0;
        check(ocpg_check_bulk(ctx_, &st), "ocpg_check_bulk");
        return st;
    }
    /* The step's changes OCPControl::CalNextTimeStep reads (Bulk::CalMaxChange
     * and the wells' BHP change). */
    ocpg_max_change_t MaxChange()
    {
        ocpg_max_change_t r{};
        check(ocpg_max_change(ctx_, &r), "ocpg_max_change");
        return r;
    }

    /* IsoT_FIM::UpdateProperty after the caller's CheckP/CheckNi: CalFlash,
     * CalKrPc, CalRock, the wells' CalTrans/CalFlux and CalRes on the
     * device state (upload the Newton update first).  Returns the norms. */
    ocpg_resnorm UpdateProperty(double dt)
    {
        check(ocpg_flash_fim(ctx_, nullptr), "ocpg_flash_fim");
        check(ocpg_calkrpc_rock(ctx_), "ocpg_calkrpc_rock");
        ocpg_resnorm r{};
        check(ocpg_calres(ctx_, dt, &r), "ocpg_calres");
        return r;
    }

    /* IsoT_FIM::AssembleMat (bulk and well blocks, rhs = -residual). */
    void AssembleMat(double dt) { check(ocpg_assemble(ctx_, dt), "ocpg_assemble"); }

    /* UpdateProperty followed by AssembleMat in one call: the same results,
     * with the flash's slowest cells overlapped with the other cells'
     * properties, residual and Jacobian.  Returns the norms. */
    ocpg_resnorm UpdatePropertyAssemble(double dt)
    {
        ocpg_resnorm r{};
        check(ocpg_update_assemble(ctx_, dt, &r, nullptr), "ocpg_update_assemble");
        return r;
    }

    /* IsoT_FIM::GetSolution on the device state from the linear solver's
     * solution u (LinearSystem's layout: bulk rows, then open wells; host
     * memory); dSmaxlim is OCPControl's NRdSmax.  Updates P, Ni, S, x_ij and
     * the wells' BHP; returns Bulk::NRdPmax / NRdNmax. */
    ocpg_nr_update GetSolution(const double* u, double dSmaxlim)
    {
        ocpg_nr_update r{};
        check(ocpg_get_solution(ctx_, u, 0, dSmaxlim, &r), "ocpg_get_solution");
        return r;
    }

    /* After the UpdateProperty that follows GetSolution: Bulk::CalNRdSmax and
     * PassFlashValue's maxNRdSSP, as IsoT_FIM::FinishNR reads them. */
    ocpg_nr_change_t NRChange()
    {
        ocpg_nr_change_t r{};
        check(ocpg_nr_change(ctx_, &r), "ocpg_nr_change");
        return r;
    }

    /* The system as LinearSystem::AssembleMatLinearSolver hands it to FASP. */
    void GetSystem(BsrSystem& s) const
    {
        s.rows = ocpg_bsr_rows(ctx_);
        s.nnzb = ocpg_bsr_nnzb(ctx_);
        s.blockDim = eos_.nc + 2;
        s.IA.resize(s.rows + 1);
        s.JA.resize(s.nnzb);
        s.val.resize((size_t)s.nnzb * s.blockDim * s.blockDim);
        s.rhs.resize((size_t)s.rows * s.blockDim);
        check(ocpg_get_bsr(ctx_, s.IA.data(), s.JA.data(), s.val.data(), s.rhs.data()), "ocpg_get_bsr");
    }

    /* IsoT_FIM::UpdateLastTimeStep / ResetToLastTimeStep. */
    void UpdateLastTimeStep() { check(ocpg_update_last_step(ctx_), "ocpg_update_last_step"); }
    void ResetToLastTimeStep() { check(ocpg_reset_to_last_step(ctx_), "ocpg_reset_to_last_step"); }

    /* Host <-> device state in the reference's layout (NULL fields skipped). */
    void Upload(const ocpg_bulk& bk) { check(ocpg_upload_bulk(ctx_, &bk), "ocpg_upload_bulk"); }
    void Download(ocpg_bulk& bk) { check(ocpg_download_bulk(ctx_, &bk), "ocpg_download_bulk"); }

    /* Well::CalFlux / CalProdWeight results of the last UpdateProperty. */
    void GetWellFlux(double* perfQi, double* wellQi, double* prodWeight) const
    {
        check(ocpg_get_well_flux(ctx_, perfQi, wellQi, prodWeight), "ocpg_get_well_flux");
    }

    const ocpg_eos& Eos() const { return eos_; }
    ocpg_ctx* Context() { return ctx_; }

private:
    ocpg_eos eos_{};
    ocpg_ctx* ctx_ = nullptr;
};

}  // namespace ocpg

#endif
