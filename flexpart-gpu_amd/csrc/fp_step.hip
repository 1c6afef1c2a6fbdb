// fp_step.hip - the fused per-particle timestep for gfx950.
//
// One kernel replaces the reference's four production dispatches
// (advection_dual_wind.wgsl, langevin_fused.wgsl, dry_deposition.wgsl,
// wet_deposition.wgsl; step order of particle_step.wgsl:684-883).  Each
// particle's hot state (44 B) is read once from SoA arrays with coalesced
// loads, advanced through advection, Hanna turbulence, Philox/Box-Muller
// noise, the Langevin update with vertical sub-steps and PBL reflection, dry
// and wet deposition, and written back once (56 B with the velocity memory).
// Optionally the concentration grid is accumulated in the same pass.
//
// The kernel is bound by vector-ALU issue, not by HBM: its arithmetic is laid
// out so that a wavefront whose particles sit in different turbulence
// regimes (neutral, unstable, stable Hanna profiles) or Langevin branches
// still evaluates each expensive primitive - division, exp, the binary64
// log/exp behind pow, sqrt - once, with regime-selected operands, instead of
// once per divergent branch.  Every selected operation is the same operation
// on the same operands as in the reference formula, so values are unchanged.
//
// Floating point: compiled with -ffp-contract=off and correctly rounded
// divide/sqrt; transcendentals come from fp_math.h.  With the operation order
// kept as in the reference, results are reproducible bit for bit.
#include "fp_physics.hpp"

namespace fp {

// Words of a record that the step writes for an active particle.
__device__ __forceinline__ bool stepped_field(uint32_t f, uint32_t n_species)
{
    return f <= F_FLAGS || (f >= F_VEL_U && f <= F_TURB_W) || f == F_CBT ||
           (f >= F_MASS0 && f < F_MASS0 + n_species);
}

// A word of a particle set: with SOA32 (a set of at most kSoa32Slots slots)
// addressed with a 32-bit offset from the set's base, else from the field's
// own base.
#if defined(__HIP_DEVICE_COMPILE__)
template <bool SOA32, class T>
__device__ __forceinline__ __attribute__((address_space(1))) T &soa_word(const Soa &s, uint32_t field, uint32_t slot)
{
    if constexpr (SOA32) return s.w32<T>(field, slot);
    else return s.w<T>(field, slot);
}
#else
template <bool SOA32, class T> __host__ __device__ T &soa_word(const Soa &s, uint32_t field, uint32_t slot)
{
    return s.w<T>(field, slot);
}
#endif

// One block's chunk of the step: slots first + r*BLOCK + threadIdx.x for
// r < a.ppt, temporal weight alpha and Philox base counter ctr.  With GATHER,
// slot i of a.s receives the step of slot a.gather_map[i] of a.src, taken at
// slot i (kGatherToSlot: the record was moved there by a compaction before the
// step) or at its source slot (kGatherFromSlot: the record is moved there by a
// compaction after the step) - the slot that selects its Philox counters and
// per-slot forcing.
template <uint32_t BLOCK, int NSUB, int GATHER = kGatherNone, uint32_t ROWS = 0, bool SOA32 = false>
__device__ __forceinline__ void step_chunk(const StepArgs &a, uint64_t first, float alpha, U4 ctr,
                                           const float *lh, const float *oh, GridLds &gt)
{
    const uint32_t nsub = NSUB >= 0 ? (uint32_t)NSUB : a.n_substeps;
    const Soa s = a.s;
    // Philox blocks per particle: 2 when the vertical loop takes more than two
    // normals (known at compile time for NSUB >= 0, so both blocks interleave)
    const bool two_blocks = NSUB >= 0 ? NSUB > 2 : a.blocks_per_particle > 1u;
    for (uint32_t r = 0; r < (ROWS ? ROWS : a.ppt); ++r) {
        // slot indices fit 32 bits (slot ranges are limited to 2^30 slots)
        const uint32_t i = (uint32_t)first + r * BLOCK + threadIdx.x;
        const bool inb = i < a.n;
        // the slot the state is read from
        uint32_t g = i;
        if (GATHER != kGatherNone && inb && a.gather_map) g = a.gather_map[i];
        const Soa sl = GATHER != kGatherNone ? a.src : s;
        // the slot the particle occupies during the step
        const uint32_t at = GATHER == kGatherFromSlot ? g : i;
        // The slot's whole hot state is requested at once (44 B, coalesced);
        // slots whose flags say inactive are not used further.
        uint32_t flags = 0;
        int32_t cx = 0, cy = 0, cbt = 0;
        float px = 0.0f, py = 0.0f, pz = 0.0f, tu = 0.0f, tv = 0.0f, tw = 0.0f;
        float mass[FP_MAX_SPECIES];
        if (inb) {
            flags = soa_word<SOA32, uint32_t>(sl, F_FLAGS, (uint32_t)g);
            cx = soa_word<SOA32, int32_t>(sl, F_CELL_X, (uint32_t)g); cy = soa_word<SOA32, int32_t>(sl, F_CELL_Y, (uint32_t)g);
            px = soa_word<SOA32, float>(sl, F_POS_X, (uint32_t)g); py = soa_word<SOA32, float>(sl, F_POS_Y, (uint32_t)g); pz = soa_word<SOA32, float>(sl, F_POS_Z, (uint32_t)g);
            tu = soa_word<SOA32, float>(sl, F_TURB_U, (uint32_t)g); tv = soa_word<SOA32, float>(sl, F_TURB_V, (uint32_t)g); tw = soa_word<SOA32, float>(sl, F_TURB_W, (uint32_t)g);
            cbt = soa_word<SOA32, int32_t>(sl, F_CBT, (uint32_t)g);
#pragma unroll
            for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                if (q < a.n_species) mass[q] = soa_word<SOA32, float>(sl, F_MASS0 + q, (uint32_t)g);
        }

        const bool live = inb && (flags & FLAG_ACTIVE);
        bool grid_valid = false;
        uint32_t grid_c = 0, grid_q = 0;
        if (live) {

            // ---- advection (advection_dual_wind.wgsl:254-303)
            const float nxm = a.nxm, nym = a.nym;
            const float x0 = (float)cx + px, y0 = (float)cy + py, z0 = pz;
            const Wind s0 = sample_wind(a, lh, alpha, x0, y0, z0);
            const float u0 = s0.u + tu, v0 = s0.v + tv, w0 = s0.w;
            const float qx = clampq(x0 + (u0 * a.x_scale) * a.dt_adv, 0.0f, nxm);
            const float qy = clampq(y0 + (v0 * a.y_scale) * a.dt_adv, 0.0f, nym);
            const float qz = clampq(z0 + (w0 * a.z_scale) * a.dt_adv, 0.0f, a.zmax);
            const Wind s1 = sample_wind(a, lh, alpha, qx, qy, qz);
            const float u1 = s1.u + tu, v1 = s1.v + tv, w1 = s1.w;
            const float cu = 0.5f * (u0 + u1), cv = 0.5f * (v0 + v1), cw = 0.5f * (w0 + w1);
            const float x1 = clampq(x0 + (cu * a.x_scale) * a.dt_adv, 0.0f, nxm);
            const float y1 = clampq(y0 + (cv * a.y_scale) * a.dt_adv, 0.0f, nym);
            pz = clampq(z0 + (cw * a.z_scale) * a.dt_adv, 0.0f, a.zmax);
            split_cell(x1, a.nx, nxm, cx, px);
            split_cell(y1, a.ny, nym, cy, py);

            // ---- PBL lookup + Hanna (langevin_fused.wgsl:414-427)
            const float hx = (float)cx + px, hy = (float)cy + py;
            const uint32_t pi = (uint32_t)floorf(clampq(hx, 0.0f, nxm));
            const uint32_t pj = (uint32_t)floorf(clampq(hy, 0.0f, nym));
            const uint32_t pidx = pi * a.ny + pj;
            const float4 pbl0 = load_at(a.pbl, pidx << 5), pbl1 = load_at(a.pbl, pidx << 5, 16u);

            // ---- Philox blocks per slot (langevin_fused.wgsl:430-451) and the
            // Box-Muller normals (they depend on the slot index only; drawn here,
            // while the PBL records load, so that they are not live across the
            // wind samples)
            const uint64_t cbase = (uint64_t)at * a.blocks_per_particle;
            const U4 b0 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase), a.key0, a.key1);
            U4 b1 = U4{0u, 0u, 0u, 0u};
            if (two_blocks) b1 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase + 1u), a.key0, a.key1);
            float eu, ev, ew0, ew1, ew2, ew3;
            gauss_pair(b0.x, b0.y, eu, ev);
            gauss_pair(b0.z, b0.w, ew0, ew1);
            gauss_pair(b1.x, b1.y, ew2, ew3);
            const Turb t = hanna_full(pbl0, pbl1, pz, a.log_1em3, a.lean_div != 0u);

            // ---- horizontal OU over the full step
            if (t.lean) {
                tu = ou_horizontal<true>(tu, t.sigu, t.tlu, a.dt, eu);
                tv = ou_horizontal<true>(tv, t.sigv, t.tlv, a.dt, ev);
            } else {
                tu = ou_horizontal(tu, t.sigu, t.tlu, a.dt, eu);
                tv = ou_horizontal(tv, t.sigv, t.tlv, a.dt, ev);
            }

            // ---- vertical OU with sub-steps and reflection
            if (nsub == 0u) {
                tw = t.lean ? ou_vertical<true>(tw, t.sigw, t.dsigwdz, t.tlw, a.dt, a.rho_grad, ew0)
                            : ou_vertical(tw, t.sigw, t.dsigwdz, t.tlw, a.dt, a.rho_grad, ew0);
            } else {
                const float dts = a.dt_sub;  // a.dt / n_substeps
                const float zmin = a.min_height;
                float sigw = t.sigw, dsw = t.dsigwdz, tlw = t.tlw;
#pragma unroll
                for (uint32_t sub = 0; sub < (NSUB > 0 ? (uint32_t)NSUB : nsub); ++sub) {
                    const float eta = sub == 0u ? ew0 : (sub == 1u ? ew1 : (sub == 2u ? ew2 : ew3));
                    if (t.lean) tw = ou_vertical<true>(tw, sigw, dsw, tlw, dts, a.rho_grad, eta);
                    else tw = ou_vertical(tw, sigw, dsw, tlw, dts, a.rho_grad, eta);
                    pz = pz + tw * dts;
                    reflect(pz, tw, cbt, zmin, t.h);
                    // hanna_short_w at the new height (langevin_fused.wgsl:502-513)
                    if (sub + 1u < nsub) vertical_profile(t, pz, a.log_1em3, sigw, dsw, tlw);
                }
            }

            // ---- dry deposition (dry_deposition.wgsl:70-106)
            if (a.dry_mode != FP_DRY_OFF) {
                if (!(pz >= a.dry_zcut)) {
                    float sv = a.dry_sv;
                    if (a.dry_mode == FP_DRY_PER_SLOT) sv = dry_survival(a.dry_vd_slots[at], a.dt, a.dry_href);
#pragma unroll
                    for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                        if (q < a.n_species) mass[q] = mass[q] * sv;
                }
            }
            // ---- wet deposition (wet_deposition.wgsl:73-111)
            if (a.wet_mode != FP_WET_OFF) {
                float sv = a.wet_sv;
                if (a.wet_mode == FP_WET_PER_SLOT) {
                    sv = wet_survival(a.wet_lambda_slots[at], a.wet_fraction_slots[at], a.dt);
                } else if (a.wet_mode == FP_WET_PRECIP_FIELD) {
                    // below_cloud_scavenging_coefficient_gas (physics/wet_scavenging.rs:213-225)
                    const float pr = pos_or_zero(a.precip[pidx]);
                    float lam = 0.0f;
                    if (pr > 0.0f) lam = pos_or_zero(pos_or_zero(a.wet_a) * lds_powf(pr, pos_or_zero(a.wet_b)));
                    sv = wet_survival(lam, a.wet_fraction, a.dt);
                }
#pragma unroll
                for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                    if (q < a.n_species) mass[q] = mass[q] * sv;
            }
            if (a.loss_mode == FP_LOSS_MIN_MASS && mass[0] < a.loss_min_mass) flags &= ~FLAG_ACTIVE;

            // ---- store
            soa_word<SOA32, int32_t>(s, F_CELL_X, (uint32_t)i) = cx; soa_word<SOA32, int32_t>(s, F_CELL_Y, (uint32_t)i) = cy;
            soa_word<SOA32, float>(s, F_POS_X, (uint32_t)i) = px; soa_word<SOA32, float>(s, F_POS_Y, (uint32_t)i) = py; soa_word<SOA32, float>(s, F_POS_Z, (uint32_t)i) = pz;
            soa_word<SOA32, float>(s, F_TURB_U, (uint32_t)i) = tu; soa_word<SOA32, float>(s, F_TURB_V, (uint32_t)i) = tv; soa_word<SOA32, float>(s, F_TURB_W, (uint32_t)i) = tw;
            soa_word<SOA32, float>(s, F_VEL_U, (uint32_t)i) = cu; soa_word<SOA32, float>(s, F_VEL_V, (uint32_t)i) = cv; soa_word<SOA32, float>(s, F_VEL_W, (uint32_t)i) = cw;
            soa_word<SOA32, int32_t>(s, F_CBT, (uint32_t)i) = cbt;
            soa_word<SOA32, uint32_t>(s, F_FLAGS, (uint32_t)i) = flags;
#pragma unroll
            for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                if (q < a.n_species) soa_word<SOA32, float>(s, F_MASS0 + q, (uint32_t)i) = mass[q];

            if (a.grid_fused && (flags & FLAG_ACTIVE)) {
                grid_valid = true;
                grid_c = a.grid.fast ? grid_cell<true>(a.grid, oh, cx, cy, px, py, pz)
                                     : grid_cell(a.grid, oh, cx, cy, px, py, pz);
                grid_q = quantize_mass(mass[a.grid.species], a.grid.mass_scale);
            }
        }
        if (GATHER != kGatherNone && inb && !live) {
            // a slot the step leaves unchanged still moves its stepped words
#pragma unroll
            for (uint32_t f = 0; f < F_COUNT; ++f)
                if (stepped_field(f, a.n_species)) soa_word<SOA32, uint32_t>(s, f, (uint32_t)i) = soa_word<SOA32, uint32_t>(sl, f, (uint32_t)g);
        }
        if (a.grid_fused) grid_reduce_wave(gt, a.grid, grid_valid, grid_c, grid_q);
    }
}

template <uint32_t BLOCK>
__device__ __forceinline__ void load_tables(const StepArgs &a, float *lh, float *oh)
{
    for (uint32_t k = threadIdx.x; k < level_table_words(a.nz); k += BLOCK) lh[k] = a.level_heights[k];
    if (threadIdx.x < FP_MAX_OUTPUT_LEVELS) oh[threadIdx.x] = a.grid.outheights[threadIdx.x];
    load_math_tab<BLOCK>();
    __syncthreads();
}

// NSUB > 0 specialises the vertical sub-step count at compile time (the
// production value is 4, timeloop.rs:1111); NSUB < 0 reads it at run time.
// ROWS = 1 specialises one particle per thread.  The kernel is bound by
// vector-ALU issue and the latency of its gathers, so resident wavefronts
// matter: the single-row kernel fits 77 VGPRs (6 wavefronts per SIMD) and,
// capped at 72, runs 7 with 16 bytes of spills, which the in-place step gains
// from (1-3 %) and the gathered steps lose to.  Only blocks of up to 256
// threads (at most one wavefront per SIMD each) can fill 7 wavefronts per
// SIMD; larger blocks keep the uncapped allocation.  With a run-time row loop
// the compiler keeps loop-invariant values live across the rows and needs 128
// VGPRs (4 wavefronts).
template <uint32_t BLOCK, int NSUB, int GATHER>
constexpr int step_waves_per_eu(uint32_t rows)
{
    return rows != 1u ? 4 : (GATHER == kGatherNone && BLOCK <= 256u ? 7 : 5);
}
template <uint32_t BLOCK, int NSUB, int GATHER = kGatherNone, uint32_t ROWS = 0, bool SOA32 = false>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(step_waves_per_eu<BLOCK, NSUB, GATHER>(ROWS)))) step_kernel(StepArgs a)
{
    __shared__ float lh[kLevelTableMax];
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    __shared__ GridLds gt;
    if (a.grid_fused) grid_lds_clear<BLOCK>(gt);
    load_tables<BLOCK>(a, lh, oh);
    // the in-place step at 512 threads and more on 4 regions per XCD (measured
    // 1-2 % faster than 1 or 2 regions, 8 slower; at 256 threads 4 regions were
    // 15 % slower on the later steps); the other shapes and the gathered steps
    // keep one
    const uint32_t chunk = GATHER == kGatherNone && BLOCK >= 512u ? xcd_chunk_rounds(blockIdx.x, gridDim.x, 4u)
                                                                  : xcd_chunk(blockIdx.x, gridDim.x);
    const uint64_t first = (uint64_t)chunk * BLOCK * a.ppt;
    step_chunk<BLOCK, NSUB, GATHER, ROWS, SOA32>(a, first, a.alpha, U4{a.c0, a.c1, a.c2, a.c3}, lh, oh, gt);
    if (a.grid_fused) grid_lds_flush<BLOCK>(gt, a.grid);
}

// ---------------------------------------------------------------- ensembles
//
// All members of an ensemble advance in one launch: a block table maps each
// block to (member, first slot); every member has its own particle range,
// met bracket, Philox key/counter and output grid (EnsMember, device
// memory), and they share the step clock, so one launch (and one captured
// graph node) covers members of any size.

__global__ void ens_clock_kernel(EnsClock *c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t k = c->next_step;
    c->step = k;
    // interpolation_alpha (timeloop.rs:2086-2120): f64, then f32, clamped
    const double t = c->t0_in_bracket + (double)k * c->dt_seconds;
    const float alpha = (float)(t / c->bracket_seconds);
    c->alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
    c->next_step = k + 1u;
}

template <uint32_t BLOCK, int NSUB, uint32_t ROWS = 0, bool SOA32 = false>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(ROWS == 1 ? 5 : 4)))
ens_step_kernel(const EnsMember *__restrict__ members, const EnsBlock *__restrict__ blocks, const EnsClock *__restrict__ clock)
{
    __shared__ float lh[kLevelTableMax];
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    __shared__ GridLds gt;
    // The tables are read through restrict-qualified pointers: nothing the
    // kernel writes aliases them, so their words are scalar loads held in
    // scalar registers, as a by-value kernel argument would be (without it
    // every use re-reads member arguments with vector loads).
    // Table entries in launch order: the dispatcher deals consecutive blocks to
    // the 8 XCDs in turn, so every XCD works on every member.  (Contiguous
    // runs of the table per XCD, as the single-context step uses, give each
    // XCD its own members and an uneven finish; measured 0.7 % slower here.)
    const EnsBlock eb = blocks[blockIdx.x];
    const EnsMember &m = members[eb.member];
    if (m.a.grid_fused) grid_lds_clear<BLOCK>(gt);
    load_tables<BLOCK>(members[0].a, lh, oh);  // one vertical grid for all members
    const U4 ctr = ctr_add(m.base_ctr[0], m.base_ctr[1], m.base_ctr[2], m.base_ctr[3],
                           (uint64_t)clock->step * m.ctr_per_step);
    step_chunk<BLOCK, NSUB, kGatherNone, ROWS, SOA32>(m.a, eb.first, clock->alpha, ctr, lh, oh, gt);
    if (m.a.grid_fused) grid_lds_flush<BLOCK>(gt, m.a.grid);
}

// Standalone gridding: ROWS rows of BLOCK particles per block share one LDS table.
template <uint32_t BLOCK, uint32_t ROWS>
__global__ void __launch_bounds__(BLOCK) grid_kernel(Soa s, uint32_t n, GridArgs g)
{
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    __shared__ GridLds gt;
    grid_lds_clear<BLOCK>(gt);
    if (threadIdx.x < FP_MAX_OUTPUT_LEVELS) oh[threadIdx.x] = g.outheights[threadIdx.x];
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * BLOCK * ROWS;
    for (uint32_t r = 0; r < ROWS; ++r) {
        const uint64_t i = first + (uint64_t)r * BLOCK + threadIdx.x;
        bool valid = false;
        uint32_t cell = 0, q = 0;
        if (i < n && (s.u(F_FLAGS)[i] & FLAG_ACTIVE)) {
            valid = true;
            cell = grid_cell(g, oh, s.i(F_CELL_X)[i], s.i(F_CELL_Y)[i], s.f(F_POS_X)[i], s.f(F_POS_Y)[i],
                             s.f(F_POS_Z)[i]);
            q = quantize_mass(s.f(F_MASS0 + g.species)[i], g.mass_scale);
        }
        grid_reduce_wave(gt, g, valid, cell, q);
    }
    grid_lds_flush<BLOCK>(gt, g);
}

template <uint32_t BLOCK>
hipError_t launch_block(const StepArgs &a, uint32_t grid, hipStream_t st)
{
    // one particle per thread: the single-row variant; the production shape
    // (4 sub-steps) with 32-bit word offsets for sets of up to kSoa32Slots
    if (a.ppt == 1u) {
        if (a.n_substeps == 4u && a.s.capacity <= kSoa32Slots)
            step_kernel<BLOCK, 4, kGatherNone, 1, true><<<grid, BLOCK, 0, st>>>(a);
        else if (a.n_substeps == 4u) step_kernel<BLOCK, 4, kGatherNone, 1><<<grid, BLOCK, 0, st>>>(a);
        else step_kernel<BLOCK, -1, kGatherNone, 1><<<grid, BLOCK, 0, st>>>(a);
    } else if (a.n_substeps == 4u) {
        step_kernel<BLOCK, 4><<<grid, BLOCK, 0, st>>>(a);
    } else {
        step_kernel<BLOCK, -1><<<grid, BLOCK, 0, st>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_step(const StepArgs &a, uint32_t block, hipStream_t st)
{
    if (a.n == 0) return hipSuccess;
    const uint64_t per_block = (uint64_t)block * a.ppt;
    const uint32_t grid = (uint32_t)((a.n + per_block - 1) / per_block);
    switch (block) {
    case 64: return launch_block<64>(a, grid, st);
    case 128: return launch_block<128>(a, grid, st);
    case 256: return launch_block<256>(a, grid, st);
    case 512: return launch_block<512>(a, grid, st);
    case 1024: return launch_block<1024>(a, grid, st);
    default: return hipErrorInvalidValue;
    }
}

// A step fused with the record movement of a compaction: slot i of a.s gets
// the stepped words (stepped_field) of src slot gather_map[i] advanced by one
// step; the record's other words are not written (the caller keeps them equal
// in both sets).  at_source_slot = false: the compaction came before the step
// (the particle steps at slot i, exactly as after launch_compact), so the
// compacting loop never copies records separately; true: the compaction after
// the step is already known (a recomputation that replays its survivor map),
// so only the survivors are stepped, each at its own slot.
template <int GATHER>
hipError_t launch_gathered(const StepArgs &a, hipStream_t st)
{
    if (a.n == 0) return hipSuccess;
    if ((GATHER == kGatherFromSlot && !a.gather_map) || a.src.base == a.s.base) return hipErrorInvalidValue;
    const uint64_t per_block = 256ull * a.ppt;
    const uint32_t grid = (uint32_t)((a.n + per_block - 1) / per_block);
    if (a.ppt == 1u) {
        if (a.n_substeps == 4u && a.s.capacity <= kSoa32Slots && a.src.capacity <= kSoa32Slots)
            step_kernel<256, 4, GATHER, 1, true><<<grid, 256, 0, st>>>(a);
        else if (a.n_substeps == 4u) step_kernel<256, 4, GATHER, 1><<<grid, 256, 0, st>>>(a);
        else step_kernel<256, -1, GATHER, 1><<<grid, 256, 0, st>>>(a);
    } else if (a.n_substeps == 4u) {
        step_kernel<256, 4, GATHER><<<grid, 256, 0, st>>>(a);
    } else {
        step_kernel<256, -1, GATHER><<<grid, 256, 0, st>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_step_gathered(const StepArgs &a, bool at_source_slot, hipStream_t st)
{
    if (!a.gather_map) return hipErrorInvalidValue;
    return at_source_slot ? launch_gathered<kGatherFromSlot>(a, st) : launch_gathered<kGatherToSlot>(a, st);
}

// The step of slots [0, a.n) of a.src written to the same slots of a.s (the
// stepped words only, as in launch_step_gathered): a step and a copy of the
// state in one pass.
hipError_t launch_step_out_of_place(const StepArgs &a, hipStream_t st)
{
    StepArgs b = a;
    b.gather_map = nullptr;
    return launch_gathered<kGatherToSlot>(b, st);
}

template <uint32_t BLOCK>
hipError_t launch_ens_block(const EnsMember *members, const EnsBlock *blocks, EnsClock *clock,
                            uint32_t nblocks, uint32_t ppt, bool sub4, bool soa32, hipStream_t st)
{
    ens_clock_kernel<<<1, 64, 0, st>>>(clock);
    if (nblocks == 0) return hipGetLastError();
    // one particle per thread: the single-row variant (from 128 threads per
    // block; a 64-thread block would spill)
    if constexpr (BLOCK >= 128) {
        if (sub4 && ppt == 1u) {
            if (soa32) ens_step_kernel<BLOCK, 4, 1, true><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
            else ens_step_kernel<BLOCK, 4, 1><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
            return hipGetLastError();
        }
    }
    if (sub4) ens_step_kernel<BLOCK, 4><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
    else ens_step_kernel<BLOCK, -1><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
    return hipGetLastError();
}

hipError_t launch_ens_step(const EnsMember *members, const EnsBlock *blocks, EnsClock *clock,
                           uint32_t nblocks, uint32_t block, uint32_t ppt, bool sub4, bool soa32, hipStream_t st)
{
    switch (block) {
    case 64: return launch_ens_block<64>(members, blocks, clock, nblocks, ppt, sub4, soa32, st);
    case 128: return launch_ens_block<128>(members, blocks, clock, nblocks, ppt, sub4, soa32, st);
    case 256: return launch_ens_block<256>(members, blocks, clock, nblocks, ppt, sub4, soa32, st);
    case 512: return launch_ens_block<512>(members, blocks, clock, nblocks, ppt, sub4, soa32, st);
    case 1024: return launch_ens_block<1024>(members, blocks, clock, nblocks, ppt, sub4, soa32, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_grid(const Soa &s, uint32_t n, const GridArgs &g, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    constexpr uint32_t kRows = 8;
    const uint32_t grid = (uint32_t)(((uint64_t)n + 256u * kRows - 1u) / (256u * kRows));
    grid_kernel<256, kRows><<<grid, 256, 0, st>>>(s, n, g);
    return hipGetLastError();
}

}  // namespace fp
