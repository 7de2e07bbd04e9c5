// Launchers of the kernels in qmi_kernels.hip (all windows per launch).
#pragma once
#include <hip/hip_runtime.h>

#include "qmi_device.h"

namespace qmi {
void launch_begin(const Dev& d, hipStream_t s, int force_energy, int eval_only, int rebuild);
void launch_lists(const Dev& d, hipStream_t s, int ngrp, int nwc, int nsc);
// water-water tiles alone, and all nonbonded tile kernels (with or without
// them); `energy` / `q_energy`: the kernel instances that also evaluate the
// energy terms of windows at an energy step (false: forces only)
void launch_ww(const Dev& d, hipStream_t s, bool energy = true);
// single families, for the autotuner's timing runs
void launch_q_tiles(const Dev& d, hipStream_t s, bool q_energy = true);
void launch_lrf_build(const Dev& d, hipStream_t s);
void launch_bonded(const Dev& d, hipStream_t s);
void launch_nonbonded(const Dev& d, hipStream_t s, bool with_ww = true, bool q_energy = true);
// atom_terms false: the per-atom terms (LRF evaluation, fixed/shell and
// radial solvent restraints) are left to launch_atom_terms
void launch_other_forces(const Dev& d, hipStream_t s, bool atom_terms = true);
// the per-atom terms of an MD step into their own array (Dev::fterm)
void launch_atom_terms(const Dev& d, hipStream_t s);
// the lists and forces of an MD step on three streams (see the definition)
hipError_t launch_step_forces(const Dev& d, hipStream_t s, hipStream_t s2, hipStream_t s3, const hipEvent_t ev[6],
                              bool lists, int ngrp, int nwc, int nsc, bool q_energy);
// the end of an MD step; the gradient is completed with the per-atom terms
// of launch_atom_terms, which must have run in the same step
void launch_integrate(const Dev& d, hipStream_t s, int shake_any);
void launch_temperature(const Dev& d, hipStream_t s);
void launch_refresh32(const Dev& d, hipStream_t s);
}  // namespace qmi
