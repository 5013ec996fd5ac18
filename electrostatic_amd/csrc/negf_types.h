// negf_types.h -- plain-old-data descriptors shared by the host orchestration
// and the gfx950 kernels.
#pragma once

#define NEGF_SUB_K 16                 // sites per sub-block
#define NEGF_SUB_NS 8                 // sub-blocks per segment (max)
#define NEGF_SEG_MAX (NEGF_SUB_K * NEGF_SUB_NS)  // sites per segment (max)
#define NEGF_MAX_DECIMATION_LAYERS 128
#define NEGF_SEG_DEFAULT 128                     // default segment length
#define NEGF_MAX_SEGMENTS 8192                   // segments per chain (scan is sequential)
#define NEGF_MAX_SITES (NEGF_MAX_SEGMENTS * NEGF_SEG_MAX)

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
    double SLr, SLi;     // lead self-energies (filled by the transfer launch)
    double SRr, SRi;
    int tube;            // index into the tube table
    int out_slot;        // row of per-chain outputs (integrand samples, DOS)
};

}  // namespace negf
