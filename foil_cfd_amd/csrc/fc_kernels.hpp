// Host-callable launchers of the FoilCFD HIP kernels. Plain C++ signatures
// (raw device pointers, sizes, an opaque stream handle) so the host solver
// can be compiled by the host compiler alone. Every launcher returns the
// hipError_t of its launch as an int (0 = success).
#pragma once

#include <cstddef>
#include <cstdint>

#include "fc_layout.hpp"

namespace fc {

// Collision / boundary parameters of one fused step (the reference's
// StepParams, src/sim/lbm_core.cuh:233-241, with Cs turned into the
// precomputed Smagorinsky prefactor 18*sqrt(2)*Cs^2 exactly as the reference
// launch wrapper forms it, src/sim/lbm_core.cu:785-787).
// Step kernel variants: kStepAuto is the library's default (the one-cell-
// per-lane kernel with 64-bit population addressing, measured fastest on
// MI355X; on a factor-2 patch level the last sub-step of a coarse step also
// writes the restriction records). kStepSingle runs the same one-cell kernel
// on every sub-step and leaves the records to the separate restriction pass.
// kStepBuf runs the one-cell kernel with 32-bit buffer addressing (arrays
// under 2 GiB, the default beyond) and the separate restriction pass.
// kStepWide is kStepAuto with the 64-bit slot select of wall-adjacent cells
// that grids of 2^32 / 19 cells or more take, on any grid (for testing).
enum StepVariant : int { kStepAuto = 0, kStepSingle = 1, kStepBuf = 2, kStepWide = 3 };

struct StepArgs {
    float tau0 = 0.6f;
    float magicLambda = 3.0f / 16.0f;
    float smagPrefactor = 0.0f;
    float uInlet = 0.08f;
    int writeMacro = 0;
    int variant = kStepAuto;
    // 16-bit storages: stochastic rounding of the population stores (1) or
    // round to nearest (0); the seed of this launch's random bits.
    int stochastic = 0;
    unsigned roundSeed = 0;
};

// Device view of one lattice level (non-owning).
struct LatticeView {
    void* f = nullptr;                 // kQ * cells populations, storage format below
    int storage = kStoreF32;           // Storage enum
    const std::uint8_t* kind = nullptr;  // cells kind bytes
    Dims dims;
    int periodicX = 0, periodicY = 0;  // z is always periodic
    void* feed = nullptr;              // outlet feed planes [2][kQ][ny*nz], storage format
    float* rho = nullptr;              // optional macro outputs (cells each)
    float* u = nullptr;
    float* v = nullptr;
    float* w = nullptr;
    const std::uint16_t* uwx = nullptr;  // optional wall-model slip (half bits, cells each)
    const std::uint16_t* uwy = nullptr;
    const std::uint16_t* uwz = nullptr;
};

int launchInitEquilibrium(const LatticeView& lat, float uInit, void* stream);
int launchSpanwisePerturb(const LatticeView& lat, float amplitude, float phase1, float phase2,
                          float phase3, void* stream);
// Canonical FP32 / FP64 unpadded SoA f[q*cells + cell] <-> device layout.
int launchUploadCanonical(const LatticeView& lat, const void* canon, int canonIsF64,
                          void* stream);
int launchDownloadCanonical(const LatticeView& lat, int layout, void* canon, int canonIsF64,
                            void* stream);
// One fused stream-collide step reading layout `layout` and writing the other.
int launchStep(const LatticeView& lat, int layout, const StepArgs& args, void* stream);
// Momentum-exchange force over a sparse list of fluid cells with solid links.
// partials: device scratch of 3 * partialSlots doubles; out3: device double[3].
int launchForce(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* masks,
                int count, double* partials, int partialSlots, double* out3, void* stream);
int forcePartialSlots(int count);
// The same reduction with the partials and the result in the library's
// static device scratch (no allocation): at most kNoAllocForceSlots blocks
// of force cells (kNoAllocForceSlots * 256 cells); the result is copied into
// host3 and the stream synchronized. noAllocScratch returns the scratch
// address of the current device (looked up once per device).
inline constexpr int kNoAllocForceSlots = 8192;
inline constexpr int kMaxScratchDevices = 64;
int noAllocScratch(double** out);
int launchForceNoAlloc(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* masks, int count,
                       double* host3, void* stream);
// Global FP64 mass over Fluid cells; partials sized by massPartialSlots().
int launchMass(const LatticeView& lat, int layout, double* partials, int partialSlots,
               double* out1, void* stream);
int massPartialSlots(const Dims& d);
// NaN / runaway watchdog (strided sample), verdict written into *flag.
int launchWatchdog(const LatticeView& lat, int layout, long long sampleOffset, int flagBase,
                   int* flag, void* stream);
int launchFillFloat(float* dst, long long count, float value, void* stream);

// ---- warm restart -----------------------------------------------------------
// f = feq(rho, u) of the view's macro arrays in layout L (solid cells at
// rest) and both outlet feed-plane halves: the compact-snapshot restore.
int launchInitFromMacro(const LatticeView& lat, void* stream);
// One cell of a flag edit. kind 1: newly solid, every population set to
// f = 0. kind 2: newly fluid, f = feq(mean rho, mean u) over the neighbours
// c + e_q whose bit q is set in nbMask (the host sets a bit when that
// neighbour is Fluid under the new flags, not edited itself, inside x / y,
// and has every population addressable in the current layout); no bit set
// gives the rest state. The populations are written in layout `layout`.
struct FlagEdit {
    long long cell;
    std::uint32_t nbMask;
    std::int32_t kind;
};
int launchFlagEdits(const LatticeView& lat, int layout, const FlagEdit* edits, int count, void* stream);

// Trilinear upsample of macro fields (rho, u, v, w; cells of s each) onto
// grid d (cell centres mapped by the ratio, x / y clamped, z periodic).
int launchUpsampleMacro(const float* sRho, const float* sU, const float* sV, const float* sW, const Dims& s,
                        float* dRho, float* dU, float* dV, float* dW, const Dims& d, void* stream);

// ---- boundary-layer readouts on the macro fields ---------------------------
// One delta99 probe: column ix of the mid-span plane, from the first fluid
// row above the surface row ys up to yTop (inclusive).
struct BlStation {
    int ix, ys, yTop, pad;
};
// ue = the ray's maximum in-plane speed |(u, v)|; firstY = first row whose
// speed reaches 0.99 ue (-1 when none); uNear = u at row ys + 1.
struct BlProbe {
    float ue;
    int firstY;
    float uNear;
    int pad;
};
int launchDelta99(const float* u, const float* v, int nx, long long planeOff, const BlStation* stations,
                  int count, BlProbe* out, void* stream);
// Smallest ix in [ixStart, ixEnd - 2] whose columns ix .. ix + 2 all have a
// surface row (topY >= 0, topY + 2 < ny) and u < uRev two rows above it;
// INT_MAX in *outMin when none.
int launchSeparationOnset(const float* u, const int* topY, int nx, int ny, long long planeOff, int ixStart,
                          int ixEnd, float uRev, int* outMin, void* stream);
// (v, w) at column ix packed y + ny * z.
int launchCrossflowPlane(const float* v, const float* w, const Dims& d, int ix, float* outV, float* outW,
                         void* stream);
// Signed streamwise-vorticity sums over rows [yMin, yMax] of column ix:
// out4 = zone +, zone -, ambient +, ambient - (zone = z in [z0, z1]).
int bandPartialSlots(const Dims& d);
int launchBandCirculation(const float* v, const float* w, const Dims& d, int ix, int yMin, int yMax, int z0, int z1,
                          double* partials, int slots, double* out4, void* stream);

// ---- lattice groups (one batched step launch for several lattices) --------
inline constexpr int kGroupMaxMembers = 64;
// Size of one device record of a (step, member) pair; built on the host.
std::size_t groupEntryBytes();
// Row-segment tasks of a lattice (its share of a group launch).
int groupTasks(const LatticeView& lat);
// 1 if the lattice can join a group step (16-/32-bit storage, array under
// 2 GiB, x and y not periodic).
int groupEligible(const LatticeView& lat);
// Fills one record (groupEntryBytes() bytes at out) for a member stepping
// from `layout` with `args`; taskBase = first task of the member.
int buildGroupEntry(const LatticeView& lat, int layout, const StepArgs& args, int taskBase, void* out);
// One batched step of `members` lattices sharing `storage`: table = device
// array of their records; bounds[0..members] = task prefix sums (host).
// fromL / wm: the members' common layout parity (1: layout L) and wall-model
// setting, or -1 when they differ (per-wave dispatch).
int launchGroupStep(int storage, const void* table, int members, const int* bounds, int wavesPerBlock,
                    int slabOrder, int fromL, int wm, void* stream);

// ---- wall model -------------------------------------------------------------
// Device view of a level's wall-cell list (SoA, reference WallCellListView,
// src/sim/lbm_wallmodel.cuh:47-57); uTau is the persistent EMA state.
struct WallListView {
    const long long* cellIdx = nullptr;
    const long long* sampleIdx = nullptr;
    const float* nx = nullptr;
    const float* ny = nullptr;
    const float* nz = nullptr;
    const float* ys = nullptr;
    const std::uint32_t* mask = nullptr;
    float* uTau = nullptr;
    int count = 0;
};

inline constexpr int kMlpIn = 3;
inline constexpr int kMlpHidden = 32;

// Learned u_tau network (BF16 weight bits + FP32 biases / normalisation).
struct MlpParams {
    float inMean[4];
    float inInvStd[4];
    float envLo[4];  // training envelope in log-feature space (ln ys, ln ut, ln nu, ln Re_y)
    float envHi[4];
    float outMean, outStd;
    std::uint16_t w1[kMlpHidden * kMlpIn];
    float b1[kMlpHidden];
    std::uint16_t w2[kMlpHidden * kMlpHidden];
    float b2[kMlpHidden];
    std::uint16_t w3[kMlpHidden];
    float b3;
};

enum WallMode : int { kWallOff = 0, kWallNewton = 1, kWallMlp = 2 };

struct WallParams {
    float nuLat = 0.05f;     // level viscosity incl. the startup ramp
    float utCutoff = 1e-4f;  // 1e-3 * u_lat
    float emaAlpha = 0.25f;  // lbm_wallmodel.cuh:35
    int mode = kWallNewton;
    const MlpParams* mlp = nullptr;  // device pointer (mode kWallMlp)
};

struct WallStats {
    float sumYplus;
    int maxYplusBits;
    int activeCells;
    int clampedCells;
    int mlpCells;
};

// One wall-model update of a level (reads the post-collision state in
// `layout`, writes the half-bit slip arrays). harvest (optional): 4 floats per
// listed cell (y_s, |u_t|, nu, Newton u_tau or -1) for network training.
int launchWallUpdate(const LatticeView& lat, int layout, const WallListView& wl, std::uint16_t* uwx,
                     std::uint16_t* uwy, std::uint16_t* uwz, const WallParams& p, WallStats* stats,
                     float* harvest, void* stream);

// ---- learned u_tau network: training and batched inference (fc_mlp.hip) ---
// Master parameters: 1217 FP32 values, W1[32][3], b1[32], W2[32][32] (row =
// output unit), b2[32], w3[32], b3. X: normalised features, SoA [3][n];
// T: normalised targets [n].
inline constexpr int kMlpParamCount = 1217;
inline constexpr int kMlpTile = 32;  // samples per wavefront tile; batches are multiples
// idx[batch] = counter-based hash of (seed, step, slot) mod n.
int launchMlpBatch(int* idx, int batch, long long n, unsigned seed, unsigned step, void* stream);
// Gradient of the batch-mean squared error at P into grad[1217]; loss[0] =
// sum of squared errors. part: (batch/32) x 1217 floats, lossPart: batch/32.
int launchMlpGradient(const float* X, long long n, const float* T, const int* idx, int batch, const float* P,
                      float* part, float* lossPart, float* grad, float* loss, void* stream);
// One Adam step (step counts from 1) on the master parameters.
int launchMlpAdam(float* P, float* M, float* V, const float* grad, float lr, float beta1, float beta2, float eps,
                  int step, void* stream);
// out[i] = network u_tau of samples (4 floats each: y_s, |u_t|, nu, -), or -1
// outside the envelope.
int launchMlpInfer(const MlpParams* m, const float* samples, int count, float* out, void* stream);

// ---- tracer particles (reference src/render/particles.cuh:29-46) -----------
struct ParticleParams {
    float dtSteps = 1.0f;
    float ageRate = 0.002f;
    float inletFraction = 0.3f;
    unsigned seed = 0;
};

// pos4: 4 floats per particle (x, y, z in lattice-cell space, age in [0, 1]).
int launchParticleSeed(float* pos4, int count, const Dims& d, unsigned seed, void* stream);
// kind: the solver's kind bytes (low nibble = cell flag) or null. vel4: null,
// or device scratch of 4 floats per cell that first receives (u, v, w, 0) of
// every cell so each trilinear corner is one 16-byte read. orderScratch
// (with vel4; particleOrderScratchBytes(count, d) bytes): the particles are
// first sorted by cell and advanced in that order, so neighbouring lanes
// share the gathered lines. Every option gives the same particles.
std::size_t particleOrderScratchBytes(int count, const Dims& d);
int launchParticleAdvect(float* pos4, int count, const float* u, const float* v, const float* w,
                         const std::uint8_t* kind, const Dims& d, const ParticleParams& prm, float* vel4,
                         void* orderScratch, std::size_t orderBytes, void* stream);

// ---- q-LIBB pre-pass ----------------------------------------------------------
// Links: cell index + packed (q | byte << 8 | ff << 16); scratch: count doubles.
// ---- two-level refinement (fc_refine.hip) ---------------------------------
// Patch box in coarse cells: fine grid covers [x0,x1) x [y0,y1) x [0,nz) at
// `factor` (fine dims factor*W, factor*H, factor*nz).
struct PatchBoxI {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
};
struct RefineGeom {
    PatchBoxI box;
    int factor = 2;
    Dims coarse, fine;
};
// Coarse post-collision deviations of the ring cells (list bandXY, pairs of
// x, y) at the current time into band[(q * nz + z) * nband + i], FP32.
int launchBandSnapshot(const LatticeView& coarse, int layout, const RefineGeom& g, const int* bandXY, int nband,
                       float* band, void* stream);
// Coarse-to-fine fill (shell from band + live coarse, or full volume from the
// live coarse array alone), written in the fine level's layout fLayout.
// Shell fills of the m sub-steps of one parent step interpolate the same two
// parent time levels: with `cache` (fillCacheFloats(g) floats) cacheMode 1
// stores them, cacheMode 2 reuses them (nullptr / 0: no cache).
long long fillCacheFloats(const RefineGeom& g);
int launchFineFill(const LatticeView& coarse, int cLayout, const float* band, const int* bandIndex, int nband,
                   const LatticeView& fine, int fLayout, const RefineGeom& g, float timeWeight, float neqScale,
                   int fullVolume, float* cache, int cacheMode, void* stream);
// Fine-to-coarse restriction into the coarse array in layout cLayout;
// optional coarse macro fields (all four or none). With `scratch`
// (restrictScratchFloats(g) floats of device memory) the child averages run
// as a separate coalesced pass first; with nullptr the restriction computes
// them itself per tile halo.
long long restrictScratchFloats(const RefineGeom& g);
// recordsReady: the scratch already holds this step's child-average records
// (written by launchStepRecords), so the child-average pass is skipped.
int launchFineRestrict(const LatticeView& fine, int fLayout, const LatticeView& coarse, int cLayout,
                       const RefineGeom& g, float neqScaleInv, float* rho, float* u, float* v, float* w,
                       float* scratch, int recordsReady, void* stream);

// Fine step (as launchStep with the default variant) that also writes the
// restriction's child-average records into `scratch` (restrictScratchFloats
// floats), bit-identical to the child-average pass of launchFineRestrict.
// Factor 2 and a fine level that is not periodic in x / y only
// (stepRecordsSupported); the caller uses it for the last fine sub-step of a
// coarse step when nothing changes the fine state before the restriction.
int stepRecordsSupported(const LatticeView& fine, const RefineGeom& g);
int launchStepRecords(const LatticeView& fine, int layout, const StepArgs& args, const RefineGeom& g,
                      float* scratch, void* stream);

int launchQLibbPrepass(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* packed,
                       int count, double* scratch, void* stream);

}  // namespace fc
