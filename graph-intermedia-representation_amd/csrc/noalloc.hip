// girhip_bfs_noalloc: GIR::bfs for one source with no device allocation,
// for callers that cannot allocate (code under HIP graph capture, tight
// serving loops).  It runs the kernels of girhip_bfs's single-source path
// (direction-optimising BFS, reach.h) over buffers carved from one file-scope
// device array instead of the handle's grow-on-demand scratch.
//
// The array holds, for graphs of at most NA_VMAX vertices and NA_EMAX arcs:
// two frontier queues (int32[NA_VMAX] each), the visited / frontier / next
// bitmaps (uint64[NA_VMAX / 64] each), the level counter ring and a heavy-row
// list of NA_EMAX / HEAVY_T + 64 rows: about 36 MB per device, held from the
// library's first use on a device until the process ends.  Every call on a
// device shares that device's copy, so calls that overlap on one device race.
#include <mutex>

#include "common.h"
#include "reach.h"

namespace gir {
namespace {

constexpr int64_t NA_VMAX = 1LL << 22;
constexpr int64_t NA_EMAX = 1LL << 27;
constexpr int64_t NA_W = NA_VMAX / 64;
constexpr int64_t NA_HCAP = NA_EMAX / HEAVY_T + 64;
constexpr int64_t NA_RING = 2 * REACH_RING;

// byte offsets of the pieces (each 256-byte aligned)
constexpr int64_t al(int64_t x) { return (x + 255) & ~int64_t(255); }
constexpr int64_t O_Q0 = 0;
constexpr int64_t O_Q1 = al(O_Q0 + 4 * NA_VMAX);
constexpr int64_t O_VIS = al(O_Q1 + 4 * NA_VMAX);
constexpr int64_t O_FB = al(O_VIS + 8 * NA_W);
constexpr int64_t O_NB = al(O_FB + 8 * NA_W);
constexpr int64_t O_RING = al(O_NB + 8 * NA_W);
constexpr int64_t O_HCTL = al(O_RING + 8 * NA_RING);
constexpr int64_t O_HROW = al(O_HCTL + 8 * 2);
constexpr int64_t O_HAUX = al(O_HROW + 4 * NA_HCAP);
constexpr int64_t O_HB = al(O_HAUX + 4 * NA_HCAP);
constexpr int64_t O_HE = al(O_HB + 8 * NA_HCAP);
constexpr int64_t O_HPAY = al(O_HE + 8 * NA_HCAP);
constexpr int64_t O_HOFF = al(O_HPAY + 8 * NA_HCAP);
constexpr int64_t NA_BYTES = al(O_HOFF + 8 * (NA_HCAP + 1));

}  // namespace
int64_t outDegree(girhip_graph* h, int32_t v);  // (bfs.hip)

}  // namespace gir

static __device__ __align__(256) uint8_t scratch[gir::NA_BYTES];

namespace gir {

void* noAllocScratch(int device) {
    static std::mutex mu;
    static std::vector<void*> ptr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptr.size() <= device) ptr.resize((size_t)device + 1, nullptr);
    if (!ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            raise(GIRHIP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        ptr[device] = p;
    }
    return ptr[device];
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_bfs_noalloc(girhip_graph* h, int32_t source, int32_t* levels) {
    return guarded([&] {
        if (!h) raise(GIRHIP_EINVAL, "bfs_noalloc: null handle");
        useDevice(h);
        const int32_t V = h->V;
        if ((int64_t)V > NA_VMAX || h->E > NA_EMAX)
            raise(GIRHIP_EINVAL, "bfs_noalloc: graph exceeds the scratch array (at most 2^22 vertices and 2^27 arcs); "
                                 "use girhip_bfs");
        if (!levels || !isDevicePtr(levels)) raise(GIRHIP_EINVAL, "bfs_noalloc: levels must be device memory (V int32)");
        if (!h->na_scratch) raise(GIRHIP_ESTATE, "bfs_noalloc: scratch array not resolved");
        uint8_t* base = (uint8_t*)h->na_scratch;
        ReachBufs B;
        B.q[0] = (int32_t*)(base + O_Q0);
        B.q[1] = (int32_t*)(base + O_Q1);
        B.vis = (uint64_t*)(base + O_VIS);
        B.fbits = (uint64_t*)(base + O_FB);
        B.nbits = (uint64_t*)(base + O_NB);
        B.W = ((int64_t)V + 63) / 64;
        int64_t* ring = (int64_t*)(base + O_RING);
        const HeavyList H{(int64_t*)(base + O_HCTL), (int32_t*)(base + O_HROW), (int32_t*)(base + O_HAUX),
                          (int64_t*)(base + O_HB),   (int64_t*)(base + O_HE),   (uint64_t*)(base + O_HPAY),
                          (int64_t*)(base + O_HOFF)};
        // kernel timing would take events from a pool that may grow: off here
        const double prof = h->prm.profile;
        h->prm.profile = 0;
        struct Restore {
            girhip_graph* h;
            double p;
            ~Restore() { h->prm.profile = p; }
        } restore{h, prof};
        fillI32(h, levels, -1, V);
        if (source >= 0 && source < V) {
            // the heavy list must start empty (a previous call may have ended
            // early); the visited bitmap starts clear
            fillI64(h, H.ctl, 0, 2);
            fillU64(h, B.vis, 0, B.W);
            fillI32(h, B.q[0], source, 1);
            const int64_t m0 = outDegree(h, source);
            runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, m0, (int64_t)V, h->E, AllArcs{},
                     WriteLevel{levels}, true, ring, &H);
        }
        syncStream(h);
    });
}
