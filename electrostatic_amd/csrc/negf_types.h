// negf_types.h -- plain-old-data descriptors shared by the host orchestration
// and the gfx950 kernels.
#pragma once

#define NEGF_WAVES_PER_BLOCK 8
#define NEGF_BLOCK_THREADS 512
#define NEGF_SUB_K 16                 // sites per sub-block
#define NEGF_SUB_NS 8                 // sub-blocks per segment (max)
#define NEGF_SEG_MAX (NEGF_SUB_K * NEGF_SUB_NS)  // sites per segment (max)
#define NEGF_MAX_DECIMATION_LAYERS 128
#define NEGF_MAX_SITES (64 * NEGF_SEG_MAX)      // 64 segments of 128 sites

namespace negf {

// Kinds of per-chain output.
enum ChainKind {
    KIND_EQ = 0,     // contribution Re(coef * G_nn) per site
    KIND_NONEQ = 1,  // contribution cL*Gamma_L*|G_n0|^2 + cR*Gamma_R*|G_n,N-1|^2
    KIND_DOS = 2     // per-chain sums over sites (DOS trace and transmission)
};

// One chain = one (mode, energy) pair of one tube: a scalar tridiagonal
// problem whose links alternate t_even (link n even) and t_odd (link n odd).
struct ChainDesc {
    double Er, Ei;       // energy
    double cr, ci;       // KIND_EQ coefficient
    double cL, cR;       // KIND_NONEQ coefficients (real)
    double iL, iR;       // KIND_NONEQ integrand-sample coefficients
    double t_even;       // hopping on even links (mode beta)
    double degen;        // weight used by KIND_DOS
    double SLr, SLi;     // lead self-energies (filled by sigma_kernel)
    double SRr, SRi;
    int tube;            // index into the tube table
    int out_slot;        // row of per-chain outputs (integrand samples, DOS)
};

struct TubeDesc {
    int N;
    int seg_len;         // full segment length Lf (multiple of the sub-block K)
    int nsub;            // sub-blocks per full segment
    int pad0;
    double t_odd;        // hopping on odd links (gamma); also lead tau
    double Uc0, Uc1;     // contact potentials
    const double* U;     // device potential, N values
    long long out_off;   // offset of this tube's rows in the partial buffer
    int sample[3];       // integrand sample sites (N/2, N/4, 3N/4)
    int pad;
};

struct LaunchParams {
    ChainDesc* chains;
    const TubeDesc* tubes;
    const int* block_chain0;   // first chain of each block
    const int* block_nchain;   // chains in each block (<= 16)
    const int* block_tube;     // tube of each block
    const long long* block_part; // row of the partial buffer written by each block
    const double* Ut;          // unused (kept for layout stability)
    double* part;              // [rows][N] per-block partial sums
    double2* diagG;            // optional [chain][N]
    double* chain_out;         // optional per-chain outputs
    int* status;               // bit 0: surface-GF root failure
    int use_decimation, dec_layers, dec_max_iter, dec_rel_error;
};

}  // namespace negf
