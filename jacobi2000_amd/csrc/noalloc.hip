// The allocation-free form of j2k_cell_forward_dev: the same kernels and
// host code, with every scratch buffer carved from `scratch`, a fixed device
// array of this file, instead of the model's growable slots.
//
// `scratch` exists once per device the library runs on, for the life of the
// process. Its host-side name is only a stand-in for the symbol: the device
// address comes from hipGetSymbolAddress, looked up once per device (with
// that device current) when a cell_st model is created there, and only that
// address is handed to kernels and copies. Every call on one device shares
// the region and leaves its contents behind; nothing on this path needs it
// cleared (every buffer is written before it is read), and calls that
// overlap on one device race.
#include <algorithm>
#include <mutex>
#include <string>

#include "abi.h"
#include "cell.h"

namespace j2k {

namespace {

static __device__ __align__(256) unsigned char scratch[J2K_NOALLOC_MAX_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

unsigned char* region_of(int device) {
    if (device < 0 || device >= kMaxDevices)
        fail(J2K_ERR_LIMIT, "allocation-free scratch: device index " + std::to_string(device) + " beyond " +
                                std::to_string(kMaxDevices));
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) fail(J2K_ERR_INVALID, "allocation-free scratch: no cell_st model was created on this device");
    if (g_err[device] != hipSuccess)
        fail(J2K_ERR_NO_DEVICE, std::string("allocation-free scratch: hipGetSymbolAddress failed: ") +
                                    hipGetErrorString(g_err[device]));
    return g_ptr[device];
}

size_t aligned(size_t b) {
    if (b == 0) b = 16;
    return (b + 255) & ~(size_t)255;
}

// Bytes the forward of one request carves, slot by slot in the order
// cell_setup, the pack, cell_model and the trunk ask for them (each slot
// once; a slot asked for again at the same size is reused).
size_t bytes_needed(const Model& m, const CellTables& t) {
    const size_t M = (size_t)t.M, d = m.cfg.dim, ff = m.cfg.ff_inner;
    size_t need = aligned(t.segs.size() * sizeof(Seg)) + aligned(t.tiles.size() * sizeof(Tile)) +
                  aligned(t.slices.size() * sizeof(Slice)) + aligned(t.ctiles.size() * sizeof(CellTile)) +
                  aligned(d * sizeof(float))                       // meta vector
                  + aligned(2 * t.slices.size() * sizeof(double))  // projection targets
                  + aligned(4 * sizeof(unsigned))                  // residual words
                  + aligned(4 * M * sizeof(float))                 // packed features
                  + aligned(M * d * sizeof(float));                // residual stream
    const bool lowp = m.bf16() || m.split();
    const size_t act = m.bf16() ? 2 : 4;
    const bool rows = lowp && rows_fused_ok((int)d, (int)ff);
    need += aligned(M * d * act)        // LayerNorm output
            + aligned(3 * M * d * act)  // q, k, v
            + aligned(M * d * act)      // attention output
            + (rows ? 0 : aligned(M * ff * act));  // MLP intermediate (on chip in the row-panel kernel)
    if (m.fp8()) need += aligned(M * sizeof(float));  // the rows' maxima of each product's input
    if (m.cfg.precision == J2K_PREC_F32) {  // split-K partials of the products of few tiles
        const size_t ws = std::max({gemm_f32_workspace((int)M, (int)(3 * d), (int)d), gemm_f32_workspace((int)M, (int)d, (int)d),
                                    gemm_f32_workspace((int)M, (int)ff, (int)d), gemm_f32_workspace((int)M, (int)d, (int)ff)});
        if (ws) need += aligned(ws * sizeof(float));
    }
    return need;
}

}  // namespace

void noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_done[device] = true;
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
}

}  // namespace j2k

using namespace j2k;

extern "C" int j2k_cell_forward_dev_noalloc(void* h, const float* d_grid, const uint8_t* d_known,
                                            const float* d_meta, int T, int ny, int nx, float* d_y, void* st) {
    return guard([&] {
        if (!h) fail(J2K_ERR_INVALID, "model handle is NULL");
        if (!d_grid || !d_known || !d_meta || !d_y) fail(J2K_ERR_INVALID, "a device pointer is NULL");
        Model& m = *static_cast<Model*>(h);
        const CellDims dims{T, ny, nx};
        check_cell_dims(m, dims);
        DeviceGuard dg(m.device);
        unsigned char* base = region_of(m.device);
        CellTables t;
        t.build({dims}, attn_tokens_per_block(m.cfg.dim), false);
        const size_t need = bytes_needed(m, t);
        if (need > J2K_NOALLOC_MAX_BYTES)
            fail(J2K_ERR_OOM, "j2k_cell_forward_dev_noalloc: a " + std::to_string(T) + "x" + std::to_string(ny) + "x" +
                                  std::to_string(nx) + " grid needs " + std::to_string(need) +
                                  " bytes of scratch, the fixed region holds " +
                                  std::to_string((size_t)J2K_NOALLOC_MAX_BYTES) + " (about " +
                                  std::to_string((size_t)((double)t.M * J2K_NOALLOC_MAX_BYTES / need)) +
                                  " tokens of this model)");
        hipStream_t s = st ? static_cast<hipStream_t>(st) : m.stream;
        // Carve from the region for this call only; the model's own slots,
        // its cached tables and its graphs stay as they were. Timing is off
        // inside (its events come from a pool that may grow).
        struct Scope {
            Model& m;
            bool timing;
            Scope(Model& mm, unsigned char* b) : m(mm), timing(mm.timer.on) {
                m.timer.on = false;
                m.arena.on = true;
                m.arena.base = reinterpret_cast<char*>(b);
                m.arena.size = J2K_NOALLOC_MAX_BYTES;
                m.arena.used = 0;
                for (int i = 0; i < S_COUNT; i++) {
                    m.arena.slot[i] = nullptr;
                    m.arena.cap[i] = 0;
                }
                m.cell_tables.T = -1;  // build this grid's tables in the region
            }
            ~Scope() {
                m.arena.on = false;
                m.cell_tables.T = -1;  // the model's own table slots no longer match the cache
                m.p_tables_batch = -1;
                m.timer.on = timing;
            }
        } scope(m, base);
        CellRun run;
        cell_setup(m, s, dims, d_meta, run);
        float* x4 = m.get<float>(S_X0, 4 * (size_t)run.pack.M);
        cell_pack(s, run.pack.M, d_grid, d_known, x4);
        cell_model(m, s, run.pack, x4, d_y);
    });
}
