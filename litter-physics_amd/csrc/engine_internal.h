// Internal types shared by the MLS-MPM HIP kernels and the host engine.
//
// Device layout (HBM):
//  * particles: one f32 slab of NPF fields x capacity, structure of arrays
//    (field f of particle p at slab[f * cap + p]); two slabs alternate as the
//    committed state and the trial state.  The deformation gradient is held as
//    H = F - I and the liquid volume ratio as J - 1 so that the small elastic
//    strains of a paste near yield keep f32 precision.  Kirchhoff stress is
//    symmetric and stored as six components.  Mass, reference volume and id
//    never change during a step and live in a separate single slab.
//  * grid: dense padded node array (grid.rs:12-21) re-laid in 4x4x4-node
//    blocks of 64 consecutive nodes, so one wavefront owns one block.  Each of
//    two accumulator sets holds mass, APIC momentum sum(w (m v + m C d)) and the
//    dt-free stress sum(w V0 tau d); the momentum of the reference's P2G is
//    mom - dt (4/dx^2) force.  A per-block flag marks blocks that received mass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

enum PField {
    PX, PY, PZ,
    VX, VY, VZ,
    C00, C01, C02, C10, C11, C12, C20, C21, C22,
    H00, H01, H02, H10, H11, H12, H20, H21, H22,
    JM1,
    TXX, TYY, TZZ, TXY, TXZ, TYZ,
    EPS,
    NPF
};

enum SField { SMASS, SVOL0, SEPS, NSF };  // SEPS: plastic strain of a liquid (never changes)

enum Fail {
    F_OK = 0,
    F_INVERTED = 1,
    F_EIGEN = 2,
    F_STRETCH = 3,
    F_RETURN = 4,
    F_NONFINITE = 5,
    F_GRID = 6,
    F_DISPLACEMENT = 7,
    F_PELLET = 8,
    F_TRACTION = 9
};

constexpr int kMaxPellets = 1024;
constexpr int kMaxSpheres = 64;

// Launch-constant parameters (f32 copies of the f64 host values).
struct Params {
    float dx, inv_dx;
    int cells[3];
    int nodes[3];
    int nb[3];  // blocks per axis
    float extent[3];
    float gravity[3];
    int kind;  // 0 paste, 1 liquid
    float G, lam, ys, K, inv_n, bulk, gamma, visc;
    float face_mu[6];
};

// Accumulators of one trial step.  Reset before every attempt.
struct StepStatus {
    double wall_impulse[3];      // sum over nodes of (v' - v) m, walls and budget
    double traction_impulse[3];  // sum of signed traction deposits
    double projection;           // wall_normal_projection_energy increment
    double friction;             // wall_friction_dissipation increment
    double traction_energy;      // wall_normal_traction_energy increment
    double dissipation;          // plastic dissipation increment
    double grid_momentum[3];     // sum of m v' over nodes after walls and pellets
    unsigned int max_speed_bits;
    unsigned int max_disp_bits;
    unsigned long long fail;     // (particle << 8) | code of the first failure
    int grid_nonfinite;
    int escaped;
    int nearwall;  // particles of the trial state within half a cell of a face
    int pad;
};

// Per-pellet coupling accumulators (rigid.rs:350-363).
struct PelletAccum {
    double impulse[3];
    double angular[3];
    double grid_energy;
    double constrained;
};

// Pellet geometry for the coupling kernel (state at the start of the step), in
// f64 so that the inside test phi <= 0 on the node lattice matches the
// reference's f64 decision.
struct PelletDev {
    double pos[3];
    double axis[3];
    double vel[3];
    double omega[3];
    double radius;
    double lo[3], hi[3];
    int spheres;
    int pad;
    double offsets[kMaxSpheres];
};

struct GridPtrs {
    float* m;
    float* px;
    float* py;
    float* pz;
    float* fx;
    float* fy;
    float* fz;
    int* flag;
};

__host__ __device__ inline int node_index(const Params& P, int i, int j, int k) {
    int b = ((i >> 2) * P.nb[1] + (j >> 2)) * P.nb[2] + (k >> 2);
    return (b << 6) | ((i & 3) << 4) | ((j & 3) << 2) | (k & 3);
}

__host__ __device__ inline void node_coords(const Params& P, int index, int& i, int& j, int& k) {
    int b = index >> 6, l = index & 63;
    int bk = b % P.nb[2];
    int bj = (b / P.nb[2]) % P.nb[1];
    int bi = b / (P.nb[2] * P.nb[1]);
    i = bi * 4 + (l >> 4);
    j = bj * 4 + ((l >> 2) & 3);
    k = bk * 4 + (l & 3);
}

// Host launchers (kernels.hip).
int g2p2g_chunk();
int grid_workgroups(int nblocks);
size_t rle_temp_bytes(int n);
void launch_block_runs(hipStream_t s, void* temp, size_t temp_bytes, int n, const uint64_t* sorted_keys,
                       uint32_t* ids_tmp, uint32_t* unique, uint32_t* counts, int* nruns);
void launch_reset_status(hipStream_t s, StepStatus* st, int* nw_count);
void launch_p2g(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, const float* shared,
                GridPtrs g);
void launch_nearwall(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, int* list, int* count);
void launch_traction(hipStream_t s, const Params& P, float dt, const int* list, const int* count, int max_count,
                     const float* slab, size_t cap, const float* shared, GridPtrs g, float* trS, float* trT,
                     size_t nodes_alloc);
void launch_grid_update(hipStream_t s, const Params& P, float dt, int nblocks, GridPtrs a, float* trS, float* trT,
                        float* gv, size_t nodes_alloc, GridPtrs clear, double* partials, int* bad);
void launch_couple(hipStream_t s, const Params& P, double dx, int nblocks, GridPtrs a, float* gv, size_t nodes_alloc,
                   const PelletDev* pellets, int npellets, const int* block_start, const int* block_list,
                   PelletAccum* acc, double friction);
void launch_g2p2g(hipStream_t s, const Params& P, float dt, int n, const float* src, float* dst, size_t cap,
                  const float* shared, const float* gv, size_t nodes_alloc, GridPtrs out, const int2* work, int nwork,
                  float* pmax, double* pdiss, unsigned long long* pfail, int* pesc, int* nw_list, int* nw_count,
                  hipEvent_t start, hipEvent_t stop);
void launch_finalize(hipStream_t s, int ngrid, const double* gpart, const int* gbad, int npart, const float* pmax,
                     const double* pdiss, const unsigned long long* pfail, const int* pesc, const int* nw_count,
                     StepStatus* st);
void launch_sort_keys(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, const uint32_t* id,
                      uint64_t* keys, uint32_t* order, uint64_t* refkey, int keep_ref);
void launch_permute(hipStream_t s, int n, const uint32_t* order, const float* src, float* dst, size_t cap,
                    const float* shared_src, float* shared_dst, const uint32_t* id_src, uint32_t* id_dst,
                    const uint64_t* ref_src, uint64_t* ref_dst);
size_t sort_temp_bytes(int n, int end_bit);
void launch_radix_sort(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                       const uint32_t* vals_in, uint32_t* vals_out, int n, int end_bit);
void launch_paste_batch(hipStream_t s, const Params& P, float dt, int n, const float* h_in, float* h_out,
                        float* tau, float* inc, float* dd, int* codes);
void launch_liquid_batch(hipStream_t s, const Params& P, float dt, int n, const float* jm1, const float* c,
                         float* jm1_out, float* tau, int* codes);

}  // namespace mpm
