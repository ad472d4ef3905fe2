/*
 * dyn_nb.h — C ABI of the MI355X (gfx950) nonbonded force engine for the
 * `dynamics` crate.  Plain pointers and sizes only; every entry point returns
 * an int status (DYN_OK = 0) and never aborts; the message for the last
 * failure on a handle is available from the matching *_last_error call.
 *
 * Units follow the reference: Å, kcal/mol, kcal/(mol·Å); charges are stored
 * pre-multiplied by CHARGE_UNIT_SCALER = 18.2223 (src/non_bonded.rs:63).
 *
 * Three groups of entry points:
 *   1. dyn_nb_*      the reference's device boundary (src/gpu_interface.rs),
 *                    same inputs, outputs and semantics, over the reference's
 *                    flattened pair list; plus the neighbor-list build
 *                    (src/neighbors.rs:180-281) and the 1-4 reciprocal
 *                    correction (src/non_bonded.rs:763-790).
 *   2. dyn_pme_*     the SPME reciprocal-space object the reference takes
 *                    from the `ewald` crate (PmeRecip, get_grid_n).
 *   3. dyn_engine_*  the native path: positions + topology in, forces,
 *                    energies and virial out, with the engine owning its
 *                    cluster-pair list, exclusion masks and SPME grid.
 *
 * Host pointers unless a name ends in _device.  A handle is not thread-safe;
 * each owns one HIP stream on its device and all calls are blocking unless
 * documented otherwise.
 */
#ifndef DYN_NB_H
#define DYN_NB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DYN_OK 0
#define DYN_ERR_ARG 1       /* invalid argument (null, size, range)          */
#define DYN_ERR_HIP 2       /* HIP runtime error (message has the HIP text)  */
#define DYN_ERR_STATE 3     /* call order (e.g. compute before set_pairs)    */
#define DYN_ERR_CAPACITY 4  /* caller buffer too small; required size given  */
#define DYN_ERR_NOMEM 5     /* device allocation failed                      */
#define DYN_ERR_LIMIT 6     /* a fixed limit of the design was exceeded (a
                               cluster column or a solute row too dense);
                               calling again does not help                   */

/* Version of the library (major*10000 + minor*100 + patch). */
int dyn_version(void);
/* Number of HIP devices visible (0 when none; never fails). */
int dyn_device_count(void);
/* Message of the last failure that had no handle to hold it (create calls). */
const char* dyn_global_last_error(void);

/* ======================================================================== */
/* 1. Reference device boundary                                             */
/* ======================================================================== */

typedef struct dyn_nb_ctx dyn_nb_ctx;

typedef struct dyn_nb_opts {
  /* LJ cutoff applied by dyn_nb_compute.  The reference's CUDA kernel has
   * none (src/cuda/cuda.cu:277-279): pass INFINITY (the default when opts is
   * NULL).  Its CPU path cuts LJ at lj_cutoff (src/non_bonded.rs:1001): pass
   * that value to reproduce it. */
  float lj_cutoff;
  /* 1 = also report the energy of pairs whose target is a water site (the
   * reference drops them, src/cuda/cuda.cu:316); 0 = reference behaviour. */
  uint32_t energy_all_pairs;
} dyn_nb_opts;

/* ForcesPositsGpu::new (src/gpu_interface.rs:123-129): device buffers for
 * n_dyn solute atoms and n_water 4-site waters, Ewald cutoff and alpha. */
int dyn_nb_create(dyn_nb_ctx** out, int device, uint32_t n_dyn, uint32_t n_water,
                  float cutoff_ewald, float alpha_ewald, const dyn_nb_opts* opts);
void dyn_nb_destroy(dyn_nb_ctx* ctx);
const char* dyn_nb_last_error(const dyn_nb_ctx* ctx);

/* PerNeighborGpu::new (src/gpu_interface.rs:242-411): upload the flattened
 * pair list built by setup_pairs (src/non_bonded.rs:541-677).  Arrays have
 * n entries; index meaning as in the reference: atom_type 0 = solute, 1 =
 * water; water_site 1 = O, 2 = M, 3 = H0, 4 = H1; the u8 arrays are bools.
 * alch may be NULL (no alchemical pairs). */
int dyn_nb_set_pairs(dyn_nb_ctx* ctx, uint32_t n, const uint32_t* tgt, const uint32_t* src,
                     const float* sigma, const float* eps, const float* q_tgt,
                     const float* q_src, const uint8_t* atom_type_tgt,
                     const uint8_t* water_site_tgt, const uint8_t* atom_type_src,
                     const uint8_t* water_site_src, const uint8_t* scale14,
                     const uint8_t* calc_lj, const uint8_t* calc_coul,
                     const uint8_t* symmetric, const uint8_t* alch);

/* force_nonbonded_gpu (src/gpu_interface.rs:463-632): positions in the
 * reference's per-class flattened layout (3 floats per site), cell extent,
 * override flags and λ.  Outputs: f_dyn[3·n_dyn], f_water[12·n_water] in
 * ForcesOnWaterMol field order (f_o, f_h0, f_h1, f_m; src/solvent/mod.rs:266-273),
 * virial Σ r·F, energy (pairs whose target is a solute atom), dH/dλ.  Forces
 * are accumulated in 64-bit fixed point (2^-32 kcal/(mol·Å)), so results are
 * deterministic.  Any output pointer may be NULL. */
int dyn_nb_compute(dyn_nb_ctx* ctx, const float* pos_dyn, const float* pos_w_o,
                   const float* pos_w_m, const float* pos_w_h0, const float* pos_w_h1,
                   const float cell_extent[3], uint8_t coulomb_disabled, uint8_t lj_disabled,
                   double lambda_alch, double* f_dyn, double* f_water, double* virial,
                   double* energy, double* dh_dl);

/* Raw 64-bit fixed-point force sums of the last dyn_nb_compute, one int64
 * per component, sites ordered solute, water O, M, H0, H1 (class-major),
 * i.e. n_dyn + 4·n_water sites × 3.  For bit-exact replay checks. */
int dyn_nb_get_fixed_forces(dyn_nb_ctx* ctx, int64_t* out, uint64_t n_values);

/* The 1-4 Coulomb correction of handle_spme_recip (src/non_bonded.rs:763-790):
 * for each pair (i, j) of solute atoms adds (1/1.2 − 1)·q_i q_j/r² along r̂
 * to f_dyn[i] and its negative to f_dyn[j] (force only, no energy), and
 * r·ΔF to *virial.  f_dyn and virial are accumulated into, not overwritten. */
int dyn_nb_correct_14(dyn_nb_ctx* ctx, uint32_t n14, const uint32_t* i14, const uint32_t* j14,
                      const float* pos_dyn, const float* q_dyn, const float cell_extent[3],
                      double* f_dyn, double* virial);

/* No counterpart in the reference.  dyn_nb_correct_14 for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same arguments, kernel and results
 * (bit for bit), but no hipMalloc, hipFree or other allocating call — inputs,
 * fixed-point sums and partials are staged in a fixed per-device scratch area
 * of 6.3 MB held until the process ends.  At most 65,536 solute atoms (the
 * context's n_dyn) and 262,144 pairs per call; larger calls return
 * DYN_ERR_LIMIT.  Calls on one device share the scratch area: calls that
 * overlap there (several threads or contexts at once) race. */
int dyn_nb_correct_14_noalloc(dyn_nb_ctx* ctx, uint32_t n14, const uint32_t* i14,
                              const uint32_t* j14, const float* pos_dyn, const float* q_dyn,
                              const float cell_extent[3], double* f_dyn, double* virial);

/* build_neighbors (src/neighbors.rs:180-281): for each outer point, the inner
 * indices with |min_image(p_outer − p_inner)|² < skin_sq_w_cutoff, min image
 * and |·|² evaluated exactly as the reference's f32 CPU code does, so the
 * lists are identical to its brute-force build.  symmetric = 1 requires
 * inner == outer and returns full symmetric lists (i ≠ j) with static–static
 * pairs skipped when is_static is non-NULL.  Output is CSR: row_ptr[n_outer+1]
 * and col_idx, each row ascending (the reference's order).  If col_capacity
 * is too small the call returns DYN_ERR_CAPACITY and *n_total holds the size
 * needed. */
int dyn_nb_build_neighbors(dyn_nb_ctx* ctx, const float* posits_outer, uint32_t n_outer,
                           const float* posits_inner, uint32_t n_inner,
                           const uint8_t* is_static, const float cell_extent[3],
                           uint8_t symmetric, float skin_sq_w_cutoff, uint32_t* row_ptr,
                           uint32_t* col_idx, uint64_t col_capacity, uint64_t* n_total);

/* ======================================================================== */
/* 2. SPME reciprocal space (replaces ewald::PmeRecip)                      */
/* ======================================================================== */

typedef struct dyn_pme dyn_pme;

/* PmeRecip::new((nx,ny,nz), (lx,ly,lz), alpha) (src/non_bonded.rs:853-873);
 * order is the B-spline order (4 as the reference uses, or 5). */
int dyn_pme_create(dyn_pme** out, int device, uint32_t nx, uint32_t ny, uint32_t nz,
                   float lx, float ly, float lz, float alpha, uint32_t order);
void dyn_pme_destroy(dyn_pme* pme);
const char* dyn_pme_last_error(const dyn_pme* pme);

/* PmeRecip::forces_and_virial(&pos, &q) (src/non_bonded.rs:692-694): n
 * charges at positions wrapped into the box (3 floats each).  f_out[3n]
 * receives the reciprocal-space forces; *energy the reciprocal energy minus
 * the Ewald self energy β/√π Σq² (the convention tests/spme.rs:147-168
 * requires); *virial the analytic k-space virial (trace, same sign as Σ r·F).
 * Charge spreading accumulates in 64-bit fixed point; energy and virial are
 * reduced in FP64.  Outputs may be NULL. */
int dyn_pme_forces_and_virial(dyn_pme* pme, uint32_t n, const float* pos, const float* q,
                              float* f_out, double* energy, double* virial);

/* Fixed-point charge grid of the last call (nx·ny·nz int64, x-major, z
 * fastest), for bit-exact replay checks. */
int dyn_pme_get_fixed_grid(dyn_pme* pme, int64_t* out, uint64_t n_values);

/* get_grid_n(l, spacing): grid sizes ≥ l/spacing, rounded up to sizes whose
 * prime factors are 2, 3 and 5 (the hand-written FFT's radices). */
int dyn_pme_grid_n(float lx, float ly, float lz, float spacing, uint32_t out[3]);

/* ======================================================================== */
/* 3. Native engine                                                         */
/* ======================================================================== */

typedef struct dyn_engine dyn_engine;

#define DYN_MODE_REFERENCE 0 /* the reference's conventions (see below)      */
#define DYN_MODE_STANDARD 1  /* textbook Ewald with exclusion corrections     */

typedef struct dyn_engine_params {
  float cutoff;        /* real-space cutoff for Coulomb (and LJ), Å          */
  float lj_cutoff;     /* LJ cutoff, ≤ cutoff                                */
  float beta;          /* Ewald splitting parameter, 1/Å                     */
  uint32_t grid[3];    /* PME grid; 0 disables reciprocal space              */
  uint32_t order;      /* B-spline order, 4 or 5                             */
  uint32_t mode;       /* DYN_MODE_REFERENCE or DYN_MODE_STANDARD            */
  float list_buffer;   /* extra list radius (Å) for list reuse across calls  */
  uint32_t precision;  /* 0: FP32 arithmetic; DYN_PREC_* bits select FP64   */
  float mesh_spacing;  /* > 0: the PME grid follows the box, chosen by
                        * dyn_pme_grid_n at every set_box (the reference's
                        * regen_pme, src/non_bonded.rs:850-875); grid must
                        * then be 0.  0: the fixed grid above              */
  uint32_t item_groups; /* j groups (of 8 j slots) per work item of the pair
                         * kernel, 1..64; 0: 16.  A tuning knob: results are
                         * the same for any value to f32 summation order   */
  uint32_t single_stream; /* 1: every kernel of an evaluation on the engine's
                         * stream; 0: SPME, corrections and bonded forces on a
                         * second stream beside the list build and the pair
                         * kernel.  A tuning knob: results are bit-identical */
} dyn_engine_params;

/* Precision options (the precision sweep; the default FP32 path is the one
 * the reference's kernels run): FP64 pair arithmetic in the pair kernel
 * (no alchemical decoupling), and FP64 SPME (B-spline weights, FFT grid,
 * influence function and gather in double; the spread grid keeps its 2^-32
 * fixed point).  Forces accumulate in the same fixed point either way. */
#define DYN_PREC_PAIR_F64 1u
#define DYN_PREC_PME_F64 2u

/* Topology: n_dyn solute atoms (charges, LJ type per atom, n_types (σ, ε)
 * pairs combined by Lorentz–Berthelot), n_water OPC waters (parameters of
 * src/solvent/mod.rs:58-93), excluded 1-2/1-3 pairs and scaled 1-4 pairs as
 * (i, j) solute index pairs (2 uint32 each; setup_nonbonded_exclusion_scale_flags,
 * src/prep.rs:114-143).  Water intramolecular pairs are always excluded.
 *
 * Reference mode reproduces the reference's evaluation at a fresh neighbor
 * list: excluded pairs get no reciprocal correction, 1-4 pairs get the
 * force-only vacuum correction of src/non_bonded.rs:763-790, and
 * energy_reference counts short-range pairs that involve a solute atom
 * (src/non_bonded.rs:337-342) plus the reciprocal energy minus self energy.
 * Standard mode subtracts qᵢqⱼ·erf(βr)/r for excluded pairs (and the
 * (1−1/1.2) share for 1-4 pairs) from the reciprocal sum. */
int dyn_engine_create(dyn_engine** out, int device, const dyn_engine_params* params,
                      uint32_t n_dyn, uint32_t n_water, const float* q_dyn,
                      const int32_t* type_dyn, uint32_t n_types, const float* lj_sigma_eps,
                      uint32_t n_excl, const uint32_t* excl_pairs, uint32_t n_14,
                      const uint32_t* pairs_14);
void dyn_engine_destroy(dyn_engine* eng);
const char* dyn_engine_last_error(const dyn_engine* eng);

/* Box edge lengths (orthorhombic, origin at 0).  A new box recomputes the
 * SPME influence function on the device; with mesh_spacing, a box whose
 * grid size differs re-plans the grid. */
int dyn_engine_set_box(dyn_engine* eng, const float box[3]);

/* Positions in the reference's layout (host pointers, 3 floats per site). */
int dyn_engine_set_positions(dyn_engine* eng, const float* pos_dyn, const float* pos_w_o,
                             const float* pos_w_m, const float* pos_w_h0,
                             const float* pos_w_h1);

/* Positions already on the device: one float array of n_sites × 4 floats
 * (x, y, z, unused) with sites ordered solute, then per water O, M, H0, H1.
 * The pointer must stay valid until the next compute returns. */
int dyn_engine_set_positions_device(dyn_engine* eng, const float* xyzw_device);

#define DYN_COMPUTE_SHORT 1u     /* real-space pair forces                 */
#define DYN_COMPUTE_RECIP 2u     /* SPME reciprocal space                  */
#define DYN_COMPUTE_CORR 4u      /* exclusion and 1-4 corrections          */
#define DYN_COMPUTE_ALL 7u
#define DYN_COMPUTE_REBUILD 8u   /* force a cluster-pair list rebuild      */
#define DYN_COMPUTE_ASYNC 16u    /* enqueue only; dyn_engine_sync waits    */
#define DYN_COMPUTE_NO_ENERGY 32u /* forces only: skip energies, virial and
                                   * pair counts (dyn_engine_get_energies
                                   * then fails until a compute without it) */
/* In place of DYN_COMPUTE_RECIP: the reference's SPME cadence (SPME_RATIO,
 * src/lib.rs:1282-1341) — keep the SPME forces, energy and virial of the last
 * evaluation that ran SPME (with this box), without the 1-4 vacuum
 * correction, which accompanies a fresh SPME sum only.  Runs SPME instead
 * when nothing is kept, with a decoupled molecule (the reference evaluates
 * SPME every step then) or in DYN_MODE_STANDARD. */
#define DYN_COMPUTE_RECIP_REUSE 64u

/* Alchemical decoupling of one molecule, as the reference's
 * nonbonded_force_alchemical_kernel (src/cuda/cuda.cu:367-566) and its
 * SPME cross-term scaling (src/non_bonded.rs:709-742) with the staged
 * schedule (src/alchemical.rs:82-100).  Solute atoms [first, first + count)
 * form the decoupled molecule (the reference decouples one whole molecule,
 * src/alchemical.rs:103-108; pairs with exactly one atom in it are the
 * alchemical ones, src/non_bonded.rs:581-584, :611-612); count = 0 turns
 * decoupling off.  lambda is clamped to [0, 1].  A changed atom range
 * rebuilds the list at the next compute.  DYN_MODE_REFERENCE only.
 * dyn_energies.dh_dl returns dH/dλ. */
/* The SPME grid in use (0, 0, 0 without reciprocal space). */
int dyn_engine_get_pme_grid(dyn_engine* eng, uint32_t out[3]);

int dyn_engine_set_alchemical(dyn_engine* eng, uint32_t first, uint32_t count, double lambda);

/* One nonbonded evaluation.  Rebuilds the cluster-pair list when asked, on
 * the first call, after set_box, without a list buffer, or — decided on the
 * device by the evaluation itself — when a site has moved half the list
 * buffer since the last build; the list in use is therefore always valid,
 * also with DYN_COMPUTE_ASYNC (no host reading is involved).  A list or
 * SPME buffer that overflows is reported by the next wait as
 * DYN_ERR_CAPACITY after the buffer has been enlarged (a synchronous
 * compute evaluates once more by itself). */
int dyn_engine_compute(dyn_engine* eng, uint32_t flags);
int dyn_engine_sync(dyn_engine* eng);

/* Forces of the last compute in the reference's layout: f_dyn[3·n_dyn],
 * f_water[12·n_water] in ForcesOnWaterMol order (f_o, f_h0, f_h1, f_m). */
int dyn_engine_get_forces(dyn_engine* eng, double* f_dyn, double* f_water);
/* Raw fixed-point sums, sites in the order of set_positions_device. */
int dyn_engine_get_fixed_forces(dyn_engine* eng, int64_t* out, uint64_t n_values);
/* Device pointer to the int64 fixed-point force sums (n_sites × 3,
 * component-major: all x, then all y, then all z), valid until destroy. */
const int64_t* dyn_engine_forces_device(const dyn_engine* eng);

typedef struct dyn_energies {
  double lj;            /* LJ, all pairs (1-4 scaled)                         */
  double coul_sr;       /* erfc Coulomb, all pairs (1-4 scaled)               */
  double recip;         /* reciprocal-space energy                            */
  double self;          /* −β/√π Σq² (and net-charge term)                    */
  double excl_corr;     /* exclusion/1-4 reciprocal corrections (standard)    */
  double total;         /* lj + coul_sr + recip + self + excl_corr            */
  double reference;     /* the reference's reported energy (see create)       */
  double virial_sr;     /* Σ r·F over real-space pairs                        */
  double virial_recip;  /* k-space virial + corrections                       */
  double dh_dl;         /* dH/dλ of the alchemical decoupling (0 without)     */
  uint64_t pair_count;  /* pairs inside the cutoff that interacted, 1-4 pairs
                         * included (with DYN_COMPUTE_CORR)                   */
  uint64_t list_entries;/* cluster-pair entries in the current list           */
  uint64_t list_builds; /* list builds so far                                 */
  uint64_t pair_count_kind[3]; /* pairs of the pair kernel (pair_count less the
                                * 1-4 pairs): LJ + Coulomb, LJ only, Coulomb only */
} dyn_energies;

int dyn_engine_get_energies(dyn_engine* eng, dyn_energies* out);

/* Molecule-pair energies, as the reference's energy_between_mols
 * (src/non_bonded.rs:350-361, summed into potential_energy_between_mols at
 * :529-533; its CUDA path returns none, src/gpu_interface.rs:629).  The
 * solute is cut into n_mol molecules: mol_start[m] is the first solute atom
 * of molecule m, starting at 0 and strictly increasing (the reference's
 * checks, src/lib.rs:1522-1566); n_mol = 0 switches the matrix off.  Every
 * compute with energies then also sums the short-range pair energy (LJ +
 * erfc Coulomb, 1-4 pairs scaled) of each solute–solute pair into
 * M[m_i·n_mol + m_j] and, for m_i ≠ m_j, M[m_j·n_mol + m_i] — a pair
 * inside one molecule counts once on the diagonal.  Sums are 2⁻³² kcal/mol
 * fixed point, independent of the order of the additions. */
#define DYN_MAX_MOLECULES 4096u
int dyn_engine_set_molecules(dyn_engine* eng, uint32_t n_mol, const uint32_t* mol_start);
/* out[n_mol·n_mol] (kcal/mol, row-major) of the last compute, which must have
 * computed energies after the molecules were set. */
int dyn_engine_get_mol_energies(dyn_engine* eng, double* out, uint32_t n_mol);

/* Diagnostics (no counterpart in the reference): copies of the current
 * cluster layout and list for host replay checks.  Sizes come first: call
 * with null arrays to read n_slots, n_clusters, n_entries (groups of 8 j
 * entries) and n_items.  slot_site[n_slots] (site index or −1 for padding),
 * xyzq[4·n_slots] (wrapped x, y, z, charge), slot_type[n_slots],
 * slot_gid[n_slots] (exclusion group), cl_class[n_clusters],
 * row_ptr[n_clusters+1] (in groups), cj[8·n_entries] (j slot | class << 30;
 * padding entries name a null slot beyond the exported ones), cmask[n_entries]
 * (solute–solute groups only), items[2·n_items] (cluster, first group),
 * fx_slot[3·n_slots] (fixed-point pair forces per slot, component-major). */
int dyn_engine_debug_export(dyn_engine* eng, uint32_t* n_slots, uint32_t* n_clusters,
                            uint64_t* n_entries, int32_t* slot_site, float* xyzq,
                            int32_t* slot_type, int32_t* slot_gid, uint8_t* cl_class,
                            uint32_t* row_ptr, uint32_t* cj, uint64_t* cmask, uint32_t* items,
                            uint64_t* n_items, int64_t* fx_slot);
/* Diagnostics (no counterpart in the reference): the pair kernel's per work
 * item sums of the last compute, value-major out[9·n_items]: e_lj water,
 * e_lj solute, e_coul water, e_coul solute, virial, pair counts (solute–
 * solute, LJ, Coulomb), dH/dλ — every pair counted from both of its rows. */
int dyn_engine_debug_item_partials(dyn_engine* eng, double* out, uint64_t n_items);
/* Test hook (no counterpart in the reference): after `skip` more force
 * evaluations of this engine, the next `count` evaluations get `bits` ORed
 * into the engine's device status word before the pair kernel, as the
 * device's own checks would set them (64: a site moved half the list buffer,
 * the evaluation is void and the call repeats).  Lets tests run the error
 * and repeat paths of dyn_engine_compute, dyn_md_step and dyn_batch_step
 * without a real failure.  count = 0 switches it off. */
int dyn_engine_debug_inject_status(dyn_engine* eng, uint32_t bits, uint32_t skip, uint32_t count);
/* Diagnostics (no counterpart in the reference): uploads of kernel argument
 * packs so far in this process (packs are uploaded only when they change,
 * so steady-state steps add none). */
uint64_t dyn_debug_arg_uploads(void);
/* Diagnostics: the device-layout generation (advanced whenever a captured
 * step graph may hold stale addresses or arguments: a buffer freed, a
 * setter called, an argument copy reused; graphs of MD steps are recorded
 * only while it holds still). */
uint64_t dyn_debug_layout_generation(void);
/* Diagnostics: why the last step-graph capture was refused ("" if none was);
 * the steps then run with direct launches. */
const char* dyn_debug_capture_refusal(void);

/* Per-stage device time from HIP events recorded on the engine's stream
 * around every stage while profiling is on (enable 1), or around the pair
 * kernel only (enable 2: two events per evaluation); 0 turns it off.  Each
 * stage also reports "host:<stage>", the host time spent issuing it.
 * names/ms/counts receive up to max entries; *n the number of entries. */
int dyn_engine_profile(dyn_engine* eng, int enable);
int dyn_engine_profile_read(dyn_engine* eng, uint32_t max, const char** names, double* ms,
                            uint64_t* counts, uint32_t* n);


/* ------------------------------------------------------------------------ */
/* Device-resident MD step around an engine: positions, velocities and     */
/* forces stay on the device; dyn_md_step enqueues whole steps (integration */
/* kernels, the engine's evaluation, bonded forces) without host round      */
/* trips.  Replaces MdState::step (src/integrate.rs:62-470) for the         */
/* Langevin-middle integrator (src/integrate.rs:100-190) and, with          */
/* gamma = 0, velocity Verlet without a thermostat (:191-260): rigid OPC    */
/* water (src/solvent/opc_settle.rs:107-198) with the M-site force          */
/* projected onto O and H (src/solvent/mod.rs:397-407), SHAKE/RATTLE        */
/* constraints (src/bonded.rs:141-304), harmonic bonds and angles and       */
/* periodic dihedrals (src/bonded_forces.rs:11-200).  Units: Å, ps, amu,    */
/* kcal/mol.  Barostat, COM-motion removal and snapshots are not part of    */
/* the step.                                                                */
/* ------------------------------------------------------------------------ */
typedef struct dyn_md dyn_md;

typedef struct dyn_md_params {
  float dt;           /* ps, in (0, 0.01) */
  float temperature;  /* K, Langevin target */
  float gamma;        /* 1/ps; 0 selects velocity Verlet without thermostat */
  float shake_tol;    /* Å: SHAKE iterates until the largest correction is below it */
  uint64_t seed;      /* Langevin noise (counter-based, per step and site) */
  uint32_t spme_ratio;/* SPME every spme_ratio-th force evaluation, the kept
                       * SPME forces in between (DYN_COMPUTE_RECIP_REUSE; the
                       * reference's SPME_RATIO = 2); 0 or 1: every one */
  /* Isotropic C-rescale pressure coupling (Bernetti & Bussi, J. Chem. Phys.
   * 153, 114107 (2020)), the reference's Barostat (src/barostat.rs:324-413,
   * switched off there at :367-368; BarostatCfg defaults :28-50: 1 bar,
   * 5 ps, 4.5e-5 /bar).  Every nstpcouple-th step (0: off) the pressure
   * P = (2 K_com + W)/(3V) is measured where the reference measures it
   * (src/integrate.rs:172-190, :241-310) and the box, every pressure
   * group's centre of mass and every velocity (by 1/μ) are rescaled by
   * μ = exp(ΔlnV/3), ΔlnV = (κ/τ)(P − P₀)·nstpcouple·dt + √(2κ k_B T
   * nstpcouple·dt/(τV))·ξ clamped to ±0.1 (scale_factor, :324-351), ξ from
   * the seed's counter-based stream.  Pressure groups: each water, each
   * cluster of solute atoms joined by constraints, each other atom; K_com
   * is their centre-of-mass kinetic energy and W their virial (nonbonded +
   * bonded − Σ (r − R_group)·F), in which constraint and rigid-water forces
   * cancel.  The box is scaled about the origin (the reference scales about
   * its centre: the same up to a uniform translation).  Needs the engine's
   * list to stay valid as the box changes: every rescaling rebuilds it. */
  uint32_t nstpcouple;
  float pressure_target;  /* bar */
  float tau_p;            /* ps */
  double compressibility; /* 1/bar */
  /* How the steps of a call are enqueued (a tuning knob; results are the same
   * either way): 0 = runs of 16 steady steps are recorded into HIP graphs and
   * replayed; 1 = every kernel is launched directly. */
  uint32_t direct_launch;
  /* How list builds the device decides on are placed (a tuning knob; the
   * forces are exact either way): 0 = before the pair kernel of the
   * evaluation that needs one (when a site has moved half the list buffer);
   * 1 = staged: a build starts beside the pair kernel when the largest
   * displacement comes within max(0.25 Å, twice its recent growth per step)
   * of half the buffer, and becomes the current list at the next evaluation;
   * should a site reach half the buffer first, the call repeats from its
   * start with builds before the pair kernel. */
  uint32_t staged_builds;
} dyn_md_params;

/* The engine's box must be set.  mass_dyn[n_dyn] (amu).  Bonds: ij[2n],
 * (k kcal/mol/Å², r0 Å)[2n], U = k (r − r0)²; angles: ijk[3n] (vertex j),
 * (k kcal/mol/rad², θ0 rad)[2n], U = k (θ − θ0)²; dihedrals: ijkl[4n],
 * (k kcal/mol, periodicity, phase rad)[3n], U = k (1 + cos(nφ − phase)), the
 * periodicity an integer in [0, 6];
 * constraints: ij[2n], r0[n] (coupled constraints form clusters of at most
 * 16, as bonds to hydrogen do).  Indices are solute atoms.  The MD handle
 * drives the engine (its positions become the MD state); destroy it before
 * the engine. */
int dyn_md_create(dyn_md** out, dyn_engine* eng, const dyn_md_params* p, const float* mass_dyn,
                  uint32_t n_bonds, const uint32_t* bond_ij, const float* bond_k_r0,
                  uint32_t n_angles, const uint32_t* angle_ijk, const float* angle_k_theta0,
                  uint32_t n_dihedrals, const uint32_t* dihedral_ijkl,
                  const float* dihedral_k_n_phase, uint32_t n_constraints,
                  const uint32_t* constraint_ij, const float* constraint_r0);
void dyn_md_destroy(dyn_md* md);
const char* dyn_md_last_error(const dyn_md* md);

/* pos_dyn/vel_dyn: 3·n_dyn; pos_water: n_water × (O, M, H0, H1) × 3;
 * vel_water: n_water × (O, H0, H1) × 3. */
int dyn_md_set_state(dyn_md* md, const float* pos_dyn, const float* vel_dyn,
                     const float* pos_water, const float* vel_water);
int dyn_md_get_state(dyn_md* md, float* pos_dyn, float* vel_dyn, float* pos_water,
                     float* vel_water);

/* Runs n_steps steps and waits for them.  Force evaluations after the first
 * are enqueued without waiting; should the cluster-pair list outgrow its
 * buffers meanwhile, the engine enlarges them and the call runs the steps
 * again from the state it started from (same random stream).  Only the
 * last step's force evaluation computes nonbonded energies; the others run
 * with DYN_COMPUTE_NO_ENERGY. */
int dyn_md_step(dyn_md* md, uint32_t n_steps);

typedef struct dyn_md_energies {
  double kinetic;       /* kcal/mol over solute atoms and water O/H site velocities,
                           as the reference measures it (src/thermostat.rs:46-63) */
  double kinetic_rigid; /* kcal/mol of the rigid-body motion: solute atoms, and per
                           water its centre-of-mass and rotational kinetic energy */
  double bond, angle, dihedral;  /* kcal/mol at the last force evaluation */
  double temperature;   /* K from kinetic_rigid and dof */
  double dof;           /* 3·n_dyn − constraints + 6·n_water */
  uint64_t steps;
  double pressure;      /* bar, from the current velocities and the last force
                           evaluation's forces and virial (NaN when that
                           evaluation skipped energies) */
  double virial;        /* kcal/mol, the pressure groups' virial W */
  double kinetic_com;   /* kcal/mol, the pressure groups' centre-of-mass K */
  double volume;        /* Å³ */
  double box[3];        /* Å */
  uint64_t pcouple_steps; /* box rescalings so far */
  /* The reference's SnapshotEnergyData quantities (src/snapshot.rs:61-92),
   * from the last force evaluation (NaN when it skipped energies):
   * energy_potential_nonbonded = the engine's reported energy (its
   * `reference` field in reference mode, `total` in standard mode),
   * energy_potential_bonded = bond + angle + dihedral, energy_potential =
   * their sum, dh_dl of the decoupling, and density = total mass / volume in
   * amu/Å³ (energy_kinetic is `kinetic`, temperature, pressure and volume
   * are above; energy_potential_between_mols is dyn_engine_get_mol_energies). */
  double potential_nonbonded;
  double potential_bonded;
  double potential;
  double dh_dl;
  double density;
  uint64_t repeats;     /* step calls run again from their start (a list or
                           SPME buffer outgrew its capacity) since create */
  uint64_t graph_steps; /* steps run inside launches of captured graphs */
} dyn_md_energies;

int dyn_md_get_energies(dyn_md* md, dyn_md_energies* out);

/* ---- Batches of independent systems --------------------------------------
 * Many MD handles (each with its own engine: independent systems of uneven
 * sizes, or λ-windows of one system via dyn_engine_set_alchemical) advance
 * together: every kernel of a step is one launch for all of them, on the
 * batch's stream.  This is the batched mode the reference lacks (it steps
 * one MdState at a time, src/integrate.rs:100-317).  The handles must share
 * a device and the integrator kind (all Langevin or all plain velocity
 * Verlet); while in the batch, dyn_md_step refuses them (get/set_state,
 * get_energies and the engine queries work as before).  Each system's list
 * is rebuilt when its own sites have moved far enough (the device decides
 * per system; the build kernels of the others return at once).  Destroy the
 * batch before its members. */
typedef struct dyn_batch dyn_batch;
int dyn_batch_create(dyn_batch** out, uint32_t n, dyn_md* const* mds);
void dyn_batch_destroy(dyn_batch* b);
const char* dyn_batch_last_error(const dyn_batch* b);
/* n_steps steps of every member, then a wait (replayed from the starting
 * state after a list overflow, as dyn_md_step). */
int dyn_batch_step(dyn_batch* b, uint32_t n_steps);
/* Per-stage GPU time of the batch's steps (HIP events), as dyn_engine_profile. */
int dyn_batch_profile(dyn_batch* b, int enable);
int dyn_batch_profile_read(dyn_batch* b, uint32_t max, const char** names, double* ms,
                           uint64_t* counts, uint32_t* n);

#ifdef __cplusplus
}
#endif

#endif /* DYN_NB_H */
