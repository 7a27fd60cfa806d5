// api.hip -- the extern "C" entry points of libfiltergs_amd.so.
//
// Host orchestration of rasterizer_impl.cu:254-676 (forward / backward) and
// of the FilterGS/cuda host functions, without allocation: every buffer is
// supplied by the caller and carved here into 256-byte aligned sub-arrays.
#include <algorithm>
#include <climits>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/filtergs_amd.h"
#include "fgs_device.h"
#include "fgs_internal.h"

namespace {

thread_local std::string g_last_error;

// Blend work shape (Gaussians per LDS batch of the forward and of the
// backward), process-wide; set by fgs_set_work_shape.
int g_shape[2] = { 256, 128 };
// Forward blend walk policy (fgs_set_blend_walk).
int g_walk = 0;
// Backward tile layout (fgs_set_backward_layout): one wave per tile by default.
int g_bwd_layout = 1;
// Forward tile layout without statistics (fgs_set_forward_layout): 2 picks
// one wave per tile for launches of at least kFwdWaveMinTiles tiles.  With
// fewer tiles the single waves cannot fill the GPU: at 8160 tiles (1080p)
// mean occupancy fell from 19.6 to 11.9 waves per CU and the tiles with the
// most pairs, worked by one SIMD each, set the kernel's length.
int g_fwd_layout = 2;
// Tiles per XCD run of the batched one-wave blend kernels (fgs_set_xcd_chunk).
// 16-32 measured best on the 16 orbit views (blend backward 6.92 -> 6.80 ms).
int g_xcd_chunk = 32;
// Batched one-wave blends as persistent launches over a cost-ordered tile
// queue, runs of g_queue_chunk tiles per XCD (fgs_set_tile_queue; 0: one
// workgroup per tile in launch order).
int g_queue_chunk = 0;
int g_queue_by_cost = 1;
constexpr int64_t kFwdWaveMinTiles = 12288;
int fwd_layout_for(int64_t tiles)
{
    if (g_fwd_layout != 2) return g_fwd_layout;
    return tiles >= kFwdWaveMinTiles ? 1 : 0;
}

// Optional per-stage timing with HIP events recorded on the launch stream.
// The last two time the blend kernels alone inside their stages (the
// backward stage also clears the pair marks and runs the gradient gather).
enum Stage { ST_PREPROCESS, ST_DEPTH_SORT, ST_OFFSETS, ST_DUPLICATE, ST_TILE_SORT, ST_RANGES,
             ST_BLEND_FWD, ST_BLEND_BWD, ST_PREPROCESS_BWD, ST_BLEND_FWD_KERNEL, ST_BLEND_BWD_KERNEL,
             ST_COUNT };
bool g_prof = false;
hipEvent_t g_ev[ST_COUNT][2];
bool g_ev_ready = false;
double g_prof_ms[ST_COUNT];
uint64_t g_prof_calls[ST_COUNT];
bool g_pending[ST_COUNT];

void prof_collect()
{
    for (int k = 0; k < ST_COUNT; ++k) {
        if (!g_pending[k]) continue;
        float ms = 0.f;
        if (hipEventSynchronize(g_ev[k][1]) == hipSuccess &&
            hipEventElapsedTime(&ms, g_ev[k][0], g_ev[k][1]) == hipSuccess) {
            g_prof_ms[k] += ms;
            g_prof_calls[k] += 1;
        }
        g_pending[k] = false;
    }
}

void prof_begin(int k, hipStream_t st)
{
    if (!g_prof) return;
    if (g_pending[k]) prof_collect();
    (void)hipEventRecord(g_ev[k][0], st);
}

void prof_end(int k, hipStream_t st)
{
    if (!g_prof) return;
    (void)hipEventRecord(g_ev[k][1], st);
    g_pending[k] = true;
}

struct ProfScope {
    int k;
    hipStream_t st;
    ProfScope(int stage, hipStream_t s) : k(stage), st(s) { prof_begin(k, st); }
    ~ProfScope() { prof_end(k, st); }
};

int fail(int code, const char* fmt, const char* detail = "")
{
    char buf[512];
    snprintf(buf, sizeof(buf), fmt, detail);
    g_last_error = buf;
    return code;
}

int check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
        g_last_error = buf;
        return FGS_E_LAUNCH;
    }
    return FGS_OK;
}

int sync_check(const fgs_raster_settings* s, hipStream_t st, const char* what)
{
    if (!s->debug) return FGS_OK;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipGetLastError();
    return check(e, what);
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Sequential carving of a caller buffer.
struct Carver {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count)
    {
        T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += align256(count * sizeof(T));
        return p;
    }
};

struct Geom {
    float2* means2D;
    float* depths;
    float4* conic_opacity;
    float* rgb;
    uint32_t* tiles;
    uint4* recs;
    uint32_t* tiles_s;
    uint2* rects_s;
    uint64_t* vkeysA;
    uint64_t* vkeysB;
    uint32_t* offs;
    uint32_t* hist;
    uint32_t* partials;
    uint32_t* seg_cnt;              // batch: keys per preprocess wave
    unsigned long long* nvis;       // batch: visible (view, Gaussian) keys after the depth sort
    uint64_t* vis_masks;            // batch: visibility ballot per (view, 64-group)
    uint32_t* gid;                  // batch: Gaussian index of each element
    uint32_t* grp_tiles;            // batch: tiles per (view, 64-group)
    uint32_t* grp_base;             // batch: first gradient slot per (view, 64-group)
    uint32_t* big;                  // batch: elements of more than kGatherSolo tiles (in depths' words)
    unsigned long long* nbig;       // batch: their count
};

size_t scan_words_for_hist(size_t n)
{
    // partials needed by the scan of an n-element radix histogram table
    return fgs::scan_partials_words(fgs::radix_hist_words(n));
}

// Per-element geometry of P Gaussians, or of a batch's P = V * P64 elements
// (batch: with the visibility ballots and element -> Gaussian indices).
Geom carve_geom(void* buf, int P, size_t* total, bool batch = false)
{
    Carver c{ static_cast<char*>(buf) };
    Geom g;
    const size_t n = (size_t)(P > 0 ? P : 1);
    g.means2D = c.take<float2>(n);
    g.depths = c.take<float>(n);
    g.conic_opacity = c.take<float4>(n);
    g.rgb = c.take<float>(3 * n);
    g.tiles = c.take<uint32_t>(n);
    g.recs = c.take<uint4>(n);
    g.tiles_s = c.take<uint32_t>(n);
    g.rects_s = c.take<uint2>(n);
    g.vkeysA = c.take<uint64_t>(n);
    g.vkeysB = c.take<uint64_t>(n);
    g.offs = c.take<uint32_t>(n);
    // a batch's segmented first depth pass takes up to twice the columns
    g.hist = c.take<uint32_t>(fgs::radix_hist_words(batch ? 2 * n : n));
    size_t pw = fgs::scan_partials_words(n);
    const size_t ph = scan_words_for_hist(batch ? 2 * n : n);
    g.partials = c.take<uint32_t>(pw > ph ? pw : ph);
    g.seg_cnt = c.take<uint32_t>(n / FGS_WAVE + 1);
    g.nvis = c.take<unsigned long long>(1);
    if (batch) {
        g.vis_masks = c.take<uint64_t>(n / FGS_WAVE + 1);
        g.gid = c.take<uint32_t>(n);
        g.grp_tiles = c.take<uint32_t>(n / FGS_WAVE + 1);
        g.grp_base = c.take<uint32_t>(n / FGS_WAVE + 1);
        g.big = reinterpret_cast<uint32_t*>(g.depths);   // a batch keeps no per-element depth
        g.nbig = c.take<unsigned long long>(1);
    } else {
        g.vis_masks = nullptr;
        g.gid = nullptr;
        g.grp_tiles = nullptr;
        g.grp_base = nullptr;
        g.big = nullptr;
        g.nbig = nullptr;
    }
    if (total) *total = c.off;
    return g;
}

struct Binning {
    uint64_t* pairsA;
    uint64_t* pairsB;
    uint8_t* strips;   // per visible-list entry: bit w = it blended into strip w of its tile
    float* part;       // backward: nine gradient sums per kv pair (pre-sort position)
    uint8_t* mark;     // backward: the pair's part entry is written
    uint32_t* hist;
    uint32_t* partials;
};

Binning carve_binning(void* buf, int64_t K, size_t* total)
{
    Carver c{ static_cast<char*>(buf) };
    Binning b;
    const size_t n = (size_t)(K > 0 ? K : 1);
    b.pairsA = c.take<uint64_t>(n);
    b.pairsB = c.take<uint64_t>(n);
    b.strips = c.take<uint8_t>(n);
    b.part = c.take<float>(9 * n);
    b.mark = c.take<uint8_t>(n);
    b.hist = c.take<uint32_t>(fgs::radix_hist_words(n));
    b.partials = c.take<uint32_t>(scan_words_for_hist(n));
    if (total) *total = c.off;
    return b;
}

// A batch's kv pairs in the binning buffer: tile keys and element values as
// u32 arrays, A in the first pair array's bytes, B in the second's.
struct BatchPairs {
    uint32_t* keysA;
    uint32_t* valsA;
    uint32_t* keysB;
    uint32_t* valsB;
};

BatchPairs batch_pairs(const Binning& bn, int64_t K)
{
    const size_t n = (size_t)(K > 0 ? K : 1);
    BatchPairs p;
    p.keysA = reinterpret_cast<uint32_t*>(bn.pairsA);
    p.valsA = p.keysA + n;
    p.keysB = reinterpret_cast<uint32_t*>(bn.pairsB);
    p.valsB = p.keysB + n;
    return p;
}

struct Image {
    float* final_T;
    uint32_t* n_contrib;
    uint2* ranges;
    uint32_t* vis_last;   // per pixel: end of its blends in the tile's visible list
    uint32_t* tile_vis;   // per tile: length of the visible list
    // the tile queue of a persistent batched blend (fgs::build_tile_queue):
    // cost keys, the cost order, the radix pass's histogram and scan
    // partials, and the eight claim counters
    uint64_t* qkeysA;
    uint64_t* qkeysB;
    uint32_t* qhist;
    uint32_t* qpart;
    uint32_t* qhead;
};

// Per-pixel state of npix pixels and T tiles (a batch concatenates its views).
Image carve_image_n(void* buf, size_t npix, size_t T, size_t* total)
{
    Carver c{ static_cast<char*>(buf) };
    Image im;
    im.final_T = c.take<float>(npix > 0 ? npix : 1);
    im.n_contrib = c.take<uint32_t>(npix > 0 ? npix : 1);
    im.ranges = c.take<uint2>(T > 0 ? T : 1);
    im.vis_last = c.take<uint32_t>(npix > 0 ? npix : 1);
    im.tile_vis = c.take<uint32_t>(T > 0 ? T : 1);
    im.qkeysA = c.take<uint64_t>(T > 0 ? T : 1);
    im.qkeysB = c.take<uint64_t>(T > 0 ? T : 1);
    im.qhist = c.take<uint32_t>(fgs::radix_hist_words(T));
    im.qpart = c.take<uint32_t>(scan_words_for_hist(T));
    im.qhead = c.take<uint32_t>(8);
    if (total) *total = c.off;
    return im;
}

Image carve_image(void* buf, int W, int H, size_t* total)
{
    const size_t npix = (size_t)W * (size_t)H;
    const size_t T = (size_t)((W + 15) / 16) * (size_t)((H + 15) / 16);
    return carve_image_n(buf, npix, T, total);
}

int bit_length(uint32_t n)
{
    int b = 0;
    while (n) {
        ++b;
        n >>= 1;
    }
    return b;
}

float ogfs_log_lambda(float eps_scale, float eps_max)
{
    // forward.cu:346-351, evaluated once per launch on the host
    const float epsilon_pix = 1.0f / 255.0f;
    const float eps_base = fmaxf(0.0f, 1.0f - epsilon_pix);
    const float lambda_raw = 1.0f - (float)pow((double)eps_base, 256.0);
    float lambda_g = epsilon_pix + eps_scale * (lambda_raw - epsilon_pix);
    lambda_g = fminf(fmaxf(lambda_g, epsilon_pix), eps_max);
    return (float)log((double)lambda_g);
}

int validate_settings(const fgs_raster_settings* s)
{
    if (!s) return fail(FGS_E_INVALID, "null settings%s");
    if (s->image_width <= 0 || s->image_height <= 0)
        return fail(FGS_E_INVALID, "image size must be positive%s");
    if ((s->image_width + 15) / 16 > 65535 || (s->image_height + 15) / 16 > 65535)
        return fail(FGS_E_INVALID, "image larger than 65535 tiles per side%s");
    if (!s->viewmatrix || !s->projmatrix || !s->bg || !s->campos)
        return fail(FGS_E_INVALID, "null camera pointer%s");
    return FGS_OK;
}

int validate(const fgs_raster_settings* s, const fgs_gaussians* g)
{
    if (!s || !g) return fail(FGS_E_INVALID, "null settings or gaussians%s");
    if (g->P < 0) return fail(FGS_E_INVALID, "negative P%s");
    int rc = validate_settings(s);
    if (rc) return rc;
    if (g->P > 0 && (!g->means3D || !g->opacities))
        return fail(FGS_E_INVALID, "null means3D or opacities%s");
    if ((g->shs == nullptr) == (g->colors_precomp == nullptr) && g->P > 0)
        return fail(FGS_E_INVALID, "provide exactly one of shs / colors_precomp%s");
    if (g->P > 0 && g->shs && (g->M < (s->sh_degree + 1) * (s->sh_degree + 1) || s->sh_degree > 3 || s->sh_degree < 0))
        return fail(FGS_E_INVALID, "sh_degree must be in [0,3] with M >= (deg+1)^2%s");
    const bool sr = g->scales && g->rotations;
    if (g->P > 0 && (sr == (g->cov3D_precomp != nullptr)))
        return fail(FGS_E_INVALID, "provide exactly one of scales+rotations / cov3D_precomp%s");
    return FGS_OK;
}

fgs::PreArgs make_pre_args_settings(const fgs_raster_settings* s)
{
    fgs::PreArgs a{};
    a.D = s->sh_degree;
    a.W = s->image_width;
    a.H = s->image_height;
    a.gx = (a.W + 15) / 16;
    a.gy = (a.H + 15) / 16;
    a.tanx = s->tanfovx;
    a.tany = s->tanfovy;
    a.fy = (float)a.H / (2.0f * s->tanfovy);   // rasterizer_impl.cu:283-284
    a.fx = (float)a.W / (2.0f * s->tanfovx);
    a.mod = s->scale_modifier;
    a.log_lambda = ogfs_log_lambda(s->ogfs_epsilon_scale, s->ogfs_epsilon_max);
    a.energy_floor = s->ogfs_energy_floor;
    a.ratio_min = s->ogfs_radius_ratio_min;
    a.variant = s->variant;
    a.use_filter = s->use_filter;
    a.prefiltered = s->prefiltered;
    a.grad_mode = s->grad_mode;
    a.view = s->viewmatrix;
    a.proj = s->projmatrix;
    a.campos = s->campos;
    return a;
}

fgs::PreArgs make_pre_args(const fgs_raster_settings* s, const fgs_gaussians* g)
{
    fgs::PreArgs a = make_pre_args_settings(s);
    a.P = g->P;
    a.M = g->M;
    a.means3D = g->means3D;
    a.scales = g->cov3D_precomp ? nullptr : g->scales;
    a.rotations = g->cov3D_precomp ? nullptr : g->rotations;
    a.opacities = g->opacities;
    a.shs = g->colors_precomp ? nullptr : g->shs;
    a.colors_precomp = g->colors_precomp;
    a.cov3D_precomp = g->cov3D_precomp;
    return a;
}

__global__ void k_read_sorted(const uint64_t* pairs, const float* depths, size_t K, uint64_t* keys,
                              uint32_t* vals)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint64_t p = pairs[k];
    const uint32_t g = (uint32_t)p;
    if (keys) keys[k] = (p & 0xFFFFFFFF00000000ull) | __float_as_uint(depths[g]);
    if (vals) vals[k] = g;
}

}  // namespace

extern "C" {

int fgs_set_work_shape(int32_t fwd_batch, int32_t bwd_batch)
{
    if (fwd_batch != 128 && fwd_batch != 256)
        return fail(FGS_E_INVALID, "forward batch must be 128 or 256%s");
    if (bwd_batch != 64 && bwd_batch != 128 && bwd_batch != 256)
        return fail(FGS_E_INVALID, "backward batch must be 64, 128 or 256%s");
    g_shape[0] = fwd_batch;
    g_shape[1] = bwd_batch;
    return FGS_OK;
}

int fgs_set_blend_walk(int32_t mode)
{
    if (mode < 0 || mode > 2) return fail(FGS_E_INVALID, "walk mode must be 0, 1 or 2%s");
    g_walk = mode;
    return FGS_OK;
}

int fgs_set_backward_layout(int32_t layout)
{
    if (layout != 0 && layout != 1) return fail(FGS_E_INVALID, "backward layout must be 0 or 1%s");
    g_bwd_layout = layout;
    return FGS_OK;
}

int fgs_get_backward_layout(void) { return g_bwd_layout; }

int fgs_set_forward_layout(int32_t layout)
{
    if (layout < 0 || layout > 2) return fail(FGS_E_INVALID, "forward layout must be 0, 1 or 2%s");
    g_fwd_layout = layout;
    return FGS_OK;
}

int fgs_get_forward_layout(void) { return g_fwd_layout; }

int fgs_set_xcd_chunk(int32_t chunk)
{
    if (chunk < 0 || chunk > 4096) return fail(FGS_E_INVALID, "xcd chunk must be in [0, 4096]%s");
    g_xcd_chunk = chunk;
    return FGS_OK;
}

int fgs_set_tile_queue(int32_t chunk, int32_t by_cost)
{
    if (chunk < 0 || chunk > 4096) return fail(FGS_E_INVALID, "queue chunk must be in [0, 4096]%s");
    if (by_cost != 0 && by_cost != 1) return fail(FGS_E_INVALID, "by_cost must be 0 or 1%s");
    g_queue_chunk = chunk;
    g_queue_by_cost = by_cost;
    return FGS_OK;
}

int fgs_get_tile_queue(void) { return g_queue_chunk; }

int fgs_get_work_shape(int32_t* out2)
{
    if (!out2) return fail(FGS_E_INVALID, "null output%s");
    out2[0] = g_shape[0];
    out2[1] = g_shape[1];
    return FGS_OK;
}

int fgs_profile_enable(int32_t on)
{
    if (on && !g_ev_ready) {
        for (int k = 0; k < ST_COUNT; ++k)
            for (int j = 0; j < 2; ++j)
                if (hipEventCreate(&g_ev[k][j]) != hipSuccess)
                    return fail(FGS_E_LAUNCH, "hipEventCreate failed%s");
        g_ev_ready = true;
    }
    if (!on && g_prof) prof_collect();
    g_prof = on != 0;
    return FGS_OK;
}

int fgs_profile_read(double* ms, uint64_t* calls, int32_t n, int32_t reset)
{
    prof_collect();
    for (int k = 0; k < n && k < ST_COUNT; ++k) {
        if (ms) ms[k] = g_prof_ms[k];
        if (calls) calls[k] = g_prof_calls[k];
    }
    if (reset)
        for (int k = 0; k < ST_COUNT; ++k) {
            g_prof_ms[k] = 0.0;
            g_prof_calls[k] = 0;
        }
    return ST_COUNT;
}

const char* fgs_last_error(void) { return g_last_error.c_str(); }

int fgs_abi_version(void) { return FGS_ABI_VERSION; }
float fgs_ogfs_log_lambda(float eps_scale, float eps_max) { return ogfs_log_lambda(eps_scale, eps_max); }

size_t fgs_geom_bytes(int32_t P)
{
    size_t t = 0;
    carve_geom(nullptr, P, &t);
    return t;
}

size_t fgs_image_bytes(int32_t W, int32_t H)
{
    size_t t = 0;
    carve_image(nullptr, W, H, &t);
    return t;
}

size_t fgs_binning_bytes(int64_t K, int32_t W, int32_t H)
{
    (void)W;
    (void)H;
    size_t t = 0;
    carve_binning(nullptr, K, &t);
    return t;
}

// Phase 1 for P Gaussians described by a: preprocess (plain inputs, or
// FilterGS rows), depth sort, tile offsets; counts as fgs_forward_geometry.
static int geometry_impl(const fgs_raster_settings* s, const fgs::PreArgs& a, bool rows, void* geom,
                  int32_t* radii, uint64_t* counts, hipStream_t st)
{
    int rc;
    if (!counts || (a.P > 0 && (!geom || !radii))) return fail(FGS_E_INVALID, "null output%s");
    if ((rc = check(fgs::launch_fill_u32(counts, 0u, 8, st), "counts clear"))) return rc;
    if (a.P == 0) return sync_check(s, st, "geometry");
    Geom gm = carve_geom(geom, a.P, nullptr);
    fgs::PreOut o{};
    o.radii = radii;
    o.means2D = gm.means2D;
    o.depths = gm.depths;
    o.conic_opacity = gm.conic_opacity;
    o.rgb = gm.rgb;
    o.tiles = gm.tiles;
    o.recs = gm.recs;
    o.vkeys = gm.vkeysA;
    o.error_flag = (int*)(counts + 2);
    {
        ProfScope ps(ST_PREPROCESS, st);
        if ((rc = check(rows ? fgs::launch_preprocess_fwd_rows(a, o, st) : fgs::launch_preprocess_fwd(a, o, st),
                        "preprocess")))
            return rc;
    }
    if ((rc = sync_check(s, st, "preprocess"))) return rc;
    // depth sort of the per-Gaussian keys on bits 32..63 (4 passes -> back in A)
    {
        ProfScope ps(ST_DEPTH_SORT, st);
        if ((rc = check(fgs::radix_sort_u64(gm.vkeysA, gm.vkeysB, (size_t)a.P, 32, 4, gm.hist,
                                            gm.partials, st), "depth sort")))
            return rc;
    }
    {
        ProfScope ps(ST_OFFSETS, st);
        if ((rc = check(fgs::launch_sorted_records(gm.vkeysA, gm.recs, a.P, gm.tiles_s, gm.rects_s, st),
                        "sorted records")))
            return rc;
        if ((rc = check(fgs::launch_scan_tiles(gm.tiles_s, gm.offs, a.P, gm.partials,
                                               (unsigned long long*)counts, st),
                        "tile offsets")))
            return rc;
    }
    return sync_check(s, st, "geometry");
}

int fgs_forward_geometry(const fgs_raster_settings* s, const fgs_gaussians* g, void* geom,
                         int32_t* radii, uint64_t* counts, void* stream)
{
    int rc = validate(s, g);
    if (rc) return rc;
    return geometry_impl(s, make_pre_args(s, g), false, geom, radii, counts, (hipStream_t)stream);
}

static int validate_rows(const fgs_raster_settings* s, const fgs_filtergs_rows* f)
{
    int rc = validate_settings(s);
    if (rc) return rc;
    if (!f) return fail(FGS_E_INVALID, "null rows%s");
    if (f->N < 0 || f->n_src < 0) return fail(FGS_E_INVALID, "negative row count%s");
    if (f->N > 0 && (!f->index || !f->xyz || !f->scaling || !f->rotation || !f->opacity || !f->colors))
        return fail(FGS_E_INVALID, "null parameter array%s");
    if (f->sh_degree < 0 || f->sh_degree > 4) return fail(FGS_E_INVALID, "sh_degree must be in [0,4]%s");
    if (f->sh_degree > 0 &&
        (!f->shs || !f->campos || f->rest_coeffs < (f->sh_degree + 1) * (f->sh_degree + 1) - 1))
        return fail(FGS_E_INVALID, "sh_degree > 0 needs shs with (deg+1)^2-1 coefficients and campos%s");
    return FGS_OK;
}

int fgs_filtergs_geometry(const fgs_raster_settings* s, const fgs_filtergs_rows* f, void* geom,
                          int32_t* radii, uint64_t* counts, void* stream)
{
    int rc = validate_rows(s, f);
    if (rc) return rc;
    fgs::PreArgs a = make_pre_args_settings(s);
    a.P = f->N;
    a.means3D = f->xyz;
    a.scales = f->scaling;
    a.rotations = f->rotation;
    a.opacities = f->opacity;
    a.rows_index = f->index;
    a.rows_n_src = f->n_src;
    a.rows_dc = f->colors;
    a.rows_rest = f->shs;
    a.rows_rest_floats = 3 * f->rest_coeffs;
    a.rows_deg = f->sh_degree;
    a.rows_campos = f->campos;
    return geometry_impl(s, a, true, geom, radii, counts, (hipStream_t)stream);
}

size_t fgs_stats_words(int32_t W, int32_t H)
{
    const size_t T = (size_t)((W + 15) / 16) * (size_t)((H + 15) / 16);
    return fgs::stats_words(T);
}

int fgs_forward_render(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                       void* geom, int64_t K, void* binning, void* image, float* out_color,
                       int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                       float* kv_alpha_sum, uint32_t* kv_alpha_count, void* stream)
{
    return fgs_forward_render_stats(s, g, radii, geom, K, binning, image, out_color, out_point_id,
                                    out_weight_pixel, out_point_weight, kv_alpha_sum, kv_alpha_count,
                                    nullptr, 0.f, stream);
}

// Phase 2 for P Gaussians whose colours are colors_precomp, or the
// preprocess's when null.
static int render_impl(const fgs_raster_settings* s, int P, const float* colors_precomp, const int32_t* radii,
                void* geom, int64_t K, void* binning, void* image, float* out_color,
                int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                float* kv_alpha_sum, uint32_t* kv_alpha_count, double* stats, float kpc_used_thresh,
                hipStream_t st)
{
    int rc;
    if (K < 0 || K > 0x7FFFFFFFll) return fail(FGS_E_CAPACITY, "kv pair count out of range%s");
    if (!image || !out_color || !out_point_id || !out_weight_pixel)
        return fail(FGS_E_INVALID, "null output%s");
    if ((kv_alpha_sum == nullptr) != (kv_alpha_count == nullptr))
        return fail(FGS_E_INVALID, "kv_alpha_sum and kv_alpha_count go together%s");
    const int W = s->image_width, H = s->image_height;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16, T = gx * gy;
    Image im = carve_image(image, W, H, nullptr);
    Geom gm = carve_geom(geom, P, nullptr);
    fgs::BlendArgs b{};
    b.ranges = im.ranges;
    b.tile_vis = im.tile_vis;
    b.W = W;
    b.H = H;
    b.gx = gx;
    b.gy = gy;
    b.means2D = gm.means2D;
    b.colors = colors_precomp ? colors_precomp : gm.rgb;
    b.conic_opacity = gm.conic_opacity;
    b.bg = s->bg;
    b.batch = g_shape[0];
    b.walk = g_walk;
    b.fwd_layout = fwd_layout_for((int64_t)T);
    b.xcd_chunk = g_xcd_chunk;
    const uint64_t* sorted_pairs = nullptr;
    if (K > 0) {
        if (!binning || !geom || !radii) return fail(FGS_E_INVALID, "null binning/geometry buffer%s");
        Binning bn = carve_binning(binning, K, nullptr);
        const int passes = (bit_length((uint32_t)T) + 7) / 8;
        uint64_t* first = (passes % 2 == 0) ? bn.pairsA : bn.pairsB;
        uint64_t* second = (passes % 2 == 0) ? bn.pairsB : bn.pairsA;
        {
            ProfScope ps(ST_DUPLICATE, st);
            if ((rc = check(fgs::launch_duplicate(gm.vkeysA, gm.offs, gm.rects_s, P, gx, first, gm.recs,
                                                  st),
                            "duplicate")))
                return rc;
        }
        {
            ProfScope ps(ST_TILE_SORT, st);
            if ((rc = check(fgs::radix_sort_u64(first, second, (size_t)K, 32, passes, bn.hist,
                                                bn.partials, st), "tile sort")))
                return rc;
        }
        if ((rc = sync_check(s, st, "binning"))) return rc;
        sorted_pairs = bn.pairsA;
        b.pair_ids = reinterpret_cast<const uint32_t*>(bn.pairsA);
        b.pair_shift = 1;
        b.vis_ids = reinterpret_cast<uint32_t*>(bn.pairsB);   // free once sorted
        b.kv_strips = bn.strips;
    }
    {
        ProfScope ps(ST_RANGES, st);
        if ((rc = check(fgs::launch_ranges(sorted_pairs, (size_t)K, im.ranges, T, st), "ranges"))) return rc;
    }
    if (stats) {
        if (!(kpc_used_thresh > 0.f)) return fail(FGS_E_INVALID, "KPC used threshold must be > 0%s");
        // the u64 counters after the per-tile arrays start at zero
        double* cnt = stats + fgs::kStatsHead + 3 * (size_t)T;
        if ((rc = check(fgs::launch_fill_u32(cnt, 0u, 2 * (2 + fgs::kKpcBins), st), "stats clear")))
            return rc;
        b.stats = stats;
        b.stats_thresh = kpc_used_thresh;
    }
    {
        ProfScope ps(ST_BLEND_FWD, st);
        if ((rc = check(fgs::launch_blend_fwd(b, im.final_T, im.n_contrib, im.vis_last, out_color, out_point_id,
                                              out_weight_pixel, out_point_weight, kv_alpha_sum,
                                              kv_alpha_count, st), "blend")))
            return rc;
    }
    if (stats && (rc = check(fgs::launch_stats_finalize(stats, T, P, K, st), "stats")))
        return rc;
    return sync_check(s, st, "render");
}

int fgs_forward_render_stats(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                             void* geom, int64_t K, void* binning, void* image, float* out_color,
                             int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                             float* kv_alpha_sum, uint32_t* kv_alpha_count, double* stats,
                             float kpc_used_thresh, void* stream)
{
    int rc = validate(s, g);
    if (rc) return rc;
    return render_impl(s, g->P, g->colors_precomp, radii, geom, K, binning, image, out_color, out_point_id,
                       out_weight_pixel, out_point_weight, kv_alpha_sum, kv_alpha_count, stats,
                       kpc_used_thresh, (hipStream_t)stream);
}

int fgs_filtergs_render(const fgs_raster_settings* s, const fgs_filtergs_rows* f, const int32_t* radii,
                        void* geom, int64_t K, void* binning, void* image, float* out_color,
                        int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                        double* stats, float kpc_used_thresh, void* stream)
{
    int rc = validate_rows(s, f);
    if (rc) return rc;
    return render_impl(s, f->N, nullptr, radii, geom, K, binning, image, out_color, out_point_id,
                       out_weight_pixel, out_point_weight, nullptr, nullptr, stats, kpc_used_thresh,
                       (hipStream_t)stream);
}

int fgs_read_sorted(const void* geom, int32_t P, int64_t K, const void* binning, uint64_t* keys,
                    uint32_t* vals, void* stream)
{
    if (K <= 0) return FGS_OK;
    if (!geom || !binning) return fail(FGS_E_INVALID, "null buffer%s");
    Geom gm = carve_geom(const_cast<void*>(geom), P, nullptr);
    Binning bn = carve_binning(const_cast<void*>(binning), K, nullptr);
    hipLaunchKernelGGL(k_read_sorted, dim3((unsigned)((K + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, bn.pairsA, gm.depths, (size_t)K, keys, vals);
    return check(hipGetLastError(), "read_sorted");
}

void fgs_geom_offsets(int32_t P, size_t* out5)
{
    // carve a fictitious buffer at a non-null base and report offsets from it
    static char probe[256];
    Geom g = carve_geom(probe, P, nullptr);
    out5[0] = (size_t)((char*)g.means2D - probe);
    out5[1] = (size_t)((char*)g.depths - probe);
    out5[2] = (size_t)((char*)g.conic_opacity - probe);
    out5[3] = (size_t)((char*)g.rgb - probe);
    out5[4] = (size_t)((char*)g.tiles - probe);
}

void fgs_image_offsets(int32_t W, int32_t H, size_t* out3)
{
    static char probe[256];
    Image im = carve_image(probe, W, H, nullptr);
    out3[0] = (size_t)((char*)im.final_T - probe);
    out3[1] = (size_t)((char*)im.n_contrib - probe);
    out3[2] = (size_t)((char*)im.ranges - probe);
}

const uint32_t* fgs_image_ranges(const void* image, int32_t W, int32_t H)
{
    return reinterpret_cast<const uint32_t*>(carve_image(const_cast<void*>(image), W, H, nullptr).ranges);
}

size_t fgs_backward_ws_bytes(int32_t P)
{
    return align256((size_t)(P > 0 ? P : 1) * 3 * sizeof(float));
}

}  // extern "C"

// The backward of one view; ws holds P * 3 floats (dL/dconic), cleared here.
static int backward_impl(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                         const void* geom, int64_t K, const void* binning, const void* image,
                         const float* dL_dpix, float* dL_dmeans2D, float* dL_dcolors,
                         float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh,
                         float* dL_dscales, float* dL_drot, void* ws, void* stream)
{
    int rc = 0;
    if (!geom || !image || !radii || !dL_dpix || !dL_dmeans2D || !dL_dcolors || !dL_dopacity ||
        !dL_dmeans3D || !dL_dcov3D || !ws)
        return fail(FGS_E_INVALID, "null buffer%s");
    if (g->shs && !g->colors_precomp && !dL_dsh) return fail(FGS_E_INVALID, "dL_dsh required%s");
    if (g->scales && !g->cov3D_precomp && (!dL_dscales || !dL_drot))
        return fail(FGS_E_INVALID, "dL_dscales / dL_drot required%s");
    hipStream_t st = (hipStream_t)stream;
    const int W = s->image_width, H = s->image_height;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16;
    Geom gm = carve_geom(const_cast<void*>(geom), g->P, nullptr);
    Image im = carve_image(const_cast<void*>(image), W, H, nullptr);
    float* dconic = static_cast<float*>(ws);
    if ((rc = check(fgs::launch_fill_u32(dconic, 0u, (size_t)g->P * 3, st), "ws clear"))) return rc;
    if (K > 0) {
        if (!binning) return fail(FGS_E_INVALID, "null binning buffer%s");
        Binning bn = carve_binning(const_cast<void*>(binning), K, nullptr);
        fgs::BlendArgs b{};
        b.ranges = im.ranges;
        b.tile_vis = im.tile_vis;
        b.pair_ids = reinterpret_cast<const uint32_t*>(bn.pairsA);
        b.pair_shift = 1;
        b.vis_ids = reinterpret_cast<uint32_t*>(bn.pairsB);
        b.kv_strips = bn.strips;
        b.recs = gm.recs;
        b.part = bn.part;
        b.mark = bn.mark;
        b.part_cap = (size_t)K;
        b.W = W;
        b.H = H;
        b.gx = gx;
        b.gy = gy;
        b.means2D = gm.means2D;
        b.colors = g->colors_precomp ? g->colors_precomp : gm.rgb;
        b.conic_opacity = gm.conic_opacity;
        b.bg = s->bg;
        b.bwd_batch = g_shape[1];
        b.bwd_layout = g_bwd_layout;
        b.xcd_chunk = g_xcd_chunk;
        ProfScope ps(ST_BLEND_BWD, st);
        const fgs::GradGather gg{ gm.vkeysA, gm.tiles_s, gm.offs, (size_t)g->P, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr };
        if ((rc = check(fgs::launch_blend_bwd(b, gg, im.final_T, im.vis_last, dL_dpix, dL_dmeans2D,
                                              dconic, dL_dopacity, dL_dcolors, s->grad_mode == 1, st),
                        "blend backward")))
            return rc;
        if ((rc = sync_check(s, st, "blend backward"))) return rc;
    }
    const fgs::PreArgs a = make_pre_args(s, g);
    fgs::PreBwdIn in{ radii, dL_dmeans2D, dconic, dL_dcolors, nullptr, gm.rgb };
    fgs::PreBwdOut o{};
    o.dL_dopacity = dL_dopacity;
    o.dL_dmeans3D = dL_dmeans3D;
    o.dL_dcov3D = dL_dcov3D;
    o.dL_dsh = a.shs ? dL_dsh : nullptr;
    o.dL_dscales = a.scales ? dL_dscales : nullptr;
    o.dL_drot = a.scales ? dL_drot : nullptr;
    {
        ProfScope ps(ST_PREPROCESS_BWD, st);
        if ((rc = check(fgs::launch_preprocess_bwd(a, in, o, st), "preprocess backward"))) return rc;
    }
    return sync_check(s, st, "backward");
}

// Workspace of fgs_backward_noalloc: dL/dconic of up to
// FGS_NOALLOC_MAX_GAUSSIANS Gaussians (48 MiB), one copy on every device the
// library runs on, held until the process ends.  Host code never names it
// directly: its device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) float scratch[(size_t)FGS_NOALLOC_MAX_GAUSSIANS * 3];

static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call made with that device current.
static int scratch_address(void** out)
{
    int dev = -1;
    int rc = check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(FGS_E_INVALID, "device ordinal out of range%s");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if ((rc = check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address"))) return rc;
        if (!p) return fail(FGS_E_LAUNCH, "scratch address is null%s");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FGS_OK;
}

extern "C" {

int fgs_backward(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                 const void* geom, int64_t K, const void* binning, const void* image,
                 const float* dL_dpix, float* dL_dmeans2D, float* dL_dcolors, float* dL_dopacity,
                 float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh, float* dL_dscales,
                 float* dL_drot, void* ws, void* stream)
{
    int rc = validate(s, g);
    if (rc) return rc;
    if (g->P == 0) return FGS_OK;
    return backward_impl(s, g, radii, geom, K, binning, image, dL_dpix, dL_dmeans2D, dL_dcolors,
                         dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales, dL_drot, ws, stream);
}

int fgs_backward_noalloc(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                         const void* geom, int64_t K, const void* binning, const void* image,
                         const float* dL_dpix, float* dL_dmeans2D, float* dL_dcolors,
                         float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh,
                         float* dL_dscales, float* dL_drot, void* stream)
{
    int rc = validate(s, g);
    if (rc) return rc;
    if (g->P > FGS_NOALLOC_MAX_GAUSSIANS)
        return fail(FGS_E_CAPACITY, "fgs_backward_noalloc: more Gaussians than its workspace holds "
                                   "(FGS_NOALLOC_MAX_GAUSSIANS); use fgs_backward with a workspace%s");
    if (g->P == 0) return FGS_OK;
    void* ws = nullptr;
    if ((rc = scratch_address(&ws))) return rc;
    return backward_impl(s, g, radii, geom, K, binning, image, dL_dpix, dL_dmeans2D, dL_dcolors,
                         dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales, dL_drot, ws, stream);
}

int fgs_mark_visible(int32_t P, const float* means3D, const float* viewmatrix,
                     const float* projmatrix, uint8_t* present, void* stream)
{
    (void)projmatrix;
    if (P < 0 || (P > 0 && (!means3D || !viewmatrix || !present)))
        return fail(FGS_E_INVALID, "mark_visible: bad arguments%s");
    return check(fgs::launch_mark_visible(P, means3D, viewmatrix, present, (hipStream_t)stream),
                 "mark_visible");
}

int fgs_compute_radius(int32_t P, const float* means3D, const float* scales,
                       const float* rotations, const float* viewmatrix, const float* projmatrix,
                       float tanfovx, float tanfovy, int32_t H, int32_t W, float* radii,
                       void* stream)
{
    if (P < 0 || (P > 0 && (!means3D || !scales || !rotations || !viewmatrix || !projmatrix || !radii)))
        return fail(FGS_E_INVALID, "compute_radius: bad arguments%s");
    return check(fgs::launch_radius_aa(P, means3D, scales, rotations, viewmatrix, projmatrix, tanfovx,
                                       tanfovy, H, W, radii, (hipStream_t)stream),
                 "compute_radius");
}

int fgs_compute_radius_v2(int32_t P, const float* means3D, const float* scales,
                          const float* rotations, const float* projmatrix,
                          const float* viewmatrix, float focal_x, float focal_y, float tan_fovx,
                          float tan_fovy, float* radii, void* stream)
{
    if (P < 0 || (P > 0 && (!means3D || !scales || !rotations || !viewmatrix || !projmatrix || !radii)))
        return fail(FGS_E_INVALID, "compute_radius_v2: bad arguments%s");
    return check(fgs::launch_radius_v2(P, means3D, scales, rotations, projmatrix, viewmatrix, focal_x,
                                       focal_y, tan_fovx, tan_fovy, radii, (hipStream_t)stream),
                 "compute_radius_v2");
}

size_t fgs_lod_ws_bytes(int32_t N) { return align256((size_t)(N > 0 ? N : 1)); }

int fgs_compute_visibility_radius(int32_t N, const float* means3D, const float* log_scales,
                                  const float* rotations, const float* projmatrix,
                                  const float* viewmatrix, const int32_t* tree_node_index,
                                  const int32_t* ancestor_path, int32_t L, float focal_x,
                                  float focal_y, float tan_fovx, float tan_fovy, float padding,
                                  float radius_max, float radius_filter_threshold, uint8_t* mask,
                                  float* radius2d, float* depth, void* ws, void* stream)
{
    if (N < 0 || L < 0) return fail(FGS_E_INVALID, "compute_visibility_radius: negative size%s");
    if (N == 0) return FGS_OK;
    if (!means3D || !log_scales || !rotations || !projmatrix || !viewmatrix || !tree_node_index ||
        (L > 0 && !ancestor_path) || !mask || !radius2d || !depth || !ws)
        return fail(FGS_E_INVALID, "compute_visibility_radius: null pointer%s");
    fgs::LodArgs a{};
    a.N = N;
    a.L = L;
    a.means3d = means3D;
    a.log_scales = log_scales;
    a.rotations = rotations;
    a.proj = projmatrix;
    a.view = viewmatrix;
    a.node_index = tree_node_index;
    a.ancestor_path = ancestor_path;
    a.fx = focal_x;
    a.fy = focal_y;
    a.tanx = tan_fovx;
    a.tany = tan_fovy;
    a.padding = padding;
    a.radius_max = radius_max;
    a.radius_filter_threshold = radius_filter_threshold;
    a.mask = mask;
    a.radius2d = radius2d;
    a.depth = depth;
    a.cull_flag = static_cast<uint8_t*>(ws);
    return check(fgs::launch_lod_cut(a, (hipStream_t)stream), "compute_visibility_radius");
}

int fgs_compute_visibility_radius_parent(int32_t N, const float* means3D, const float* log_scales,
                                         const float* rotations, const float* projmatrix,
                                         const float* viewmatrix, const int32_t* tree_node_index,
                                         const int32_t* index_parent, float focal_x, float focal_y,
                                         float tan_fovx, float tan_fovy, float padding, float radius_max,
                                         float radius_filter_threshold, uint8_t* mask, float* radius2d,
                                         float* depth, void* ws, void* stream)
{
    if (N < 0) return fail(FGS_E_INVALID, "compute_visibility_radius: negative size%s");
    if (N == 0) return FGS_OK;
    if (!means3D || !log_scales || !rotations || !projmatrix || !viewmatrix || !tree_node_index ||
        !index_parent || !mask || !radius2d || !depth || !ws)
        return fail(FGS_E_INVALID, "compute_visibility_radius: null pointer%s");
    fgs::LodArgs a{};
    a.N = N;
    a.means3d = means3D;
    a.log_scales = log_scales;
    a.rotations = rotations;
    a.proj = projmatrix;
    a.view = viewmatrix;
    a.node_index = tree_node_index;
    a.index_parent = index_parent;
    a.fx = focal_x;
    a.fy = focal_y;
    a.tanx = tan_fovx;
    a.tany = tan_fovy;
    a.padding = padding;
    a.radius_max = radius_max;
    a.radius_filter_threshold = radius_filter_threshold;
    a.mask = mask;
    a.radius2d = radius2d;
    a.depth = depth;
    a.cull_flag = static_cast<uint8_t*>(ws);
    return check(fgs::launch_lod_cut(a, (hipStream_t)stream), "compute_visibility_radius");
}

int fgs_sparse_adam(int64_t n, int32_t D, const int64_t* index, const uint8_t* flag, int64_t N, float* param,
                    const float* param_sel, const float* grad, float* exp_avg, float* exp_avg_sq,
                    float* max_exp_avg_sq, double lr, double beta1, double beta2, double eps, double step,
                    int32_t* bad, void* stream)
{
    if (n < 0 || N < 0 || D <= 0) return fail(FGS_E_INVALID, "sparse_adam: bad size%s");
    if (n == 0) return FGS_OK;
    if (!index || !param || !param_sel || !grad || !exp_avg || !exp_avg_sq || !bad)
        return fail(FGS_E_INVALID, "sparse_adam: null pointer%s");
    if (!(step >= 1.0)) return fail(FGS_E_INVALID, "sparse_adam: step must be >= 1%s");
    fgs::SparseAdamArgs a{};
    a.n = n;
    a.N = N;
    a.D = D;
    a.index = index;
    a.flag = flag;
    a.param = param;
    a.param_sel = param_sel;
    a.grad = grad;
    a.exp_avg = exp_avg;
    a.exp_avg_sq = exp_avg_sq;
    a.max_exp_avg_sq = max_exp_avg_sq;
    a.beta1 = (float)beta1;
    a.beta2 = (float)beta2;
    // Python scalars of the reference (sparse_optimizer.py:51-78): 1 - beta,
    // the bias corrections and the step size in double, then used against
    // float32 tensors
    a.one_m_beta1 = (float)(1.0 - beta1);
    a.one_m_beta2 = (float)(1.0 - beta2);
    const double bc1 = 1.0 - std::pow(beta1, step);
    const double bc2 = 1.0 - std::pow(beta2, step);
    a.bc2_sqrt = (float)std::sqrt(bc2);
    a.neg_step_size = (float)(-(lr / bc1));
    a.eps = (float)eps;
    a.bad = bad;
    return check(fgs::launch_sparse_adam(a, (hipStream_t)stream), "sparse_adam");
}

int fgs_counter_update(int64_t n_vis, const int64_t* visible_index, const int32_t* radii,
                       const float* point_weight, int64_t n_pid, const int64_t* point_id,
                       const int64_t* point_count, const float* grad, int64_t N, int32_t* create_steps,
                       int16_t* visible_count, float* weights_max, float* weights_sum, int16_t* radii_max,
                       int32_t* radii_max_max, int32_t* area_sum, float* grad_sum, int32_t* bad, void* stream)
{
    if (n_vis < 0 || n_pid < 0 || N < 0) return fail(FGS_E_INVALID, "counter_update: negative size%s");
    if (n_vis == 0 && n_pid == 0) return FGS_OK;
    if (!visible_index || !radii || !point_weight || (n_pid > 0 && (!point_id || !point_count)) ||
        !create_steps || !visible_count || !weights_max || !weights_sum || !radii_max || !radii_max_max ||
        !area_sum || !grad_sum || !bad)
        return fail(FGS_E_INVALID, "counter_update: null pointer%s");
    fgs::CounterArgs a{};
    a.n_vis = n_vis;
    a.n_pid = n_pid;
    a.N = N;
    a.visible_index = visible_index;
    a.radii = radii;
    a.point_weight = point_weight;
    a.point_id = point_id;
    a.point_count = point_count;
    a.grad = grad;
    a.create_steps = create_steps;
    a.visible_count = visible_count;
    a.weights_max = weights_max;
    a.weights_sum = weights_sum;
    a.radii_max = radii_max;
    a.radii_max_max = radii_max_max;
    a.area_sum = area_sum;
    a.grad_sum = grad_sum;
    a.bad = bad;
    return check(fgs::launch_counter_update(a, (hipStream_t)stream), "counter_update");
}

int fgs_rebuild_ancestor_path(int32_t N, int32_t L, const int32_t* index_parent, const int32_t* depth,
                              int32_t* ancestor_path, int32_t* bad, void* stream)
{
    if (N < 0 || L < 0) return fail(FGS_E_INVALID, "rebuild_ancestor_path: negative size%s");
    if (N == 0 || L == 0) return FGS_OK;
    if (!index_parent || !depth || !ancestor_path || !bad)
        return fail(FGS_E_INVALID, "rebuild_ancestor_path: null pointer%s");
    return check(fgs::launch_rebuild_ancestor(N, L, index_parent, depth, ancestor_path, bad, (hipStream_t)stream),
                 "rebuild_ancestor_path");
}

namespace {
// Workspace of fgs_tree_traverse: control words, per-block counts and
// offsets, two frontiers, the id list and the per-candidate flags.
struct TravWs {
    int32_t* ctl;
    uint32_t* counts;
    uint32_t* offs;
    int32_t* fa;
    int32_t* fb;
    int32_t* out;
    uint8_t* flags;
    size_t bytes;
};
static TravWs trav_carve(void* base, int32_t N, int32_t C, int32_t n_roots)
{
    const size_t cap = (size_t)N + (size_t)n_roots;
    const size_t nflags = std::max((size_t)n_roots, cap * (size_t)C);
    Carver c{ (char*)base };
    TravWs w{};
    w.ctl = c.take<int32_t>(fgs::kTravCtl);
    w.counts = c.take<uint32_t>(2 * fgs::kTravGrid);
    w.offs = c.take<uint32_t>(2 * fgs::kTravGrid);
    w.fa = c.take<int32_t>(cap);
    w.fb = c.take<int32_t>(cap);
    w.out = c.take<int32_t>(cap);
    w.flags = c.take<uint8_t>(nflags);
    w.bytes = c.off;
    return w;
}
}  // namespace

size_t fgs_tree_traverse_ws_bytes(int32_t N, int32_t C, int32_t n_roots)
{
    if (N < 0 || C < 1 || n_roots < 0) return 0;
    return trav_carve(nullptr, N, C, n_roots).bytes;
}

int fgs_tree_traverse(int32_t N, const float* xyz, const float* scaling, const float* rotation,
                      const int32_t* node_index, const int32_t* tree, int32_t n_nodes, int32_t C,
                      const int64_t* root_index, int32_t n_roots, const float* projmatrix,
                      const float* viewmatrix, float focal_x, float focal_y, float tan_fovx,
                      float tan_fovy, float min_resolution_pixel, int32_t max_depth,
                      int32_t level_hint, void* ws, int64_t* count, void* stream)
{
    if (N < 0 || n_nodes < 0 || C < 1 || n_roots < 0)
        return fail(FGS_E_INVALID, "tree_traverse: bad size%s");
    if ((size_t)N + (size_t)n_roots > (size_t)INT32_MAX / (size_t)C)
        return fail(FGS_E_INVALID, "tree_traverse: tree too large%s");
    if (!count) return fail(FGS_E_INVALID, "tree_traverse: null count%s");
    *count = 0;
    if (n_roots == 0) return FGS_OK;
    if (!xyz || !scaling || !rotation || !node_index || (n_nodes > 0 && !tree) || !root_index ||
        !projmatrix || !viewmatrix || !ws)
        return fail(FGS_E_INVALID, "tree_traverse: null pointer%s");
    // the reference stops expanding after level min(30, max_depth)
    const int limit = std::max(0, std::min(30, (int)max_depth));
    const int levels = level_hint >= 0 ? std::min(limit, (int)level_hint) : limit;
    const bool append = levels == limit;
    TravWs w = trav_carve(ws, N, C, n_roots);
    fgs::TravArgs a{};
    a.N = N;
    a.C = C;
    a.n_nodes = n_nodes;
    a.xyz = xyz;
    a.scaling = scaling;
    a.rotation = rotation;
    a.node_index = node_index;
    a.tree = tree;
    a.proj = projmatrix;
    a.view = viewmatrix;
    a.fx = focal_x;
    a.fy = focal_y;
    a.tanx = tan_fovx;
    a.tany = tan_fovy;
    a.min_px = min_resolution_pixel;
    a.cap_frontier = N + n_roots;
    a.cap_out = N + n_roots;
    a.ctl = w.ctl;
    a.counts = w.counts;
    a.offs = w.offs;
    a.flags = w.flags;
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = check(fgs::launch_traverse(a, root_index, n_roots, w.fa, w.fb, w.out, levels, append, st),
                    "tree_traverse")))
        return rc;
    int32_t ctl[fgs::kTravCtl];
    if ((rc = check(hipMemcpyAsync(ctl, w.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st), "tree_traverse")))
        return rc;
    if ((rc = check(hipStreamSynchronize(st), "tree_traverse"))) return rc;
    if (ctl[fgs::kTravErr] & 1)
        return fail(FGS_E_INVALID, "tree_traverse: a root, node row or child id is out of range%s");
    const int64_t last = ctl[levels & 1 ? fgs::kTravFrontierB : fgs::kTravFrontierA];
    int64_t n = ctl[fgs::kTravOut] + (append ? last : 0);
    if ((ctl[fgs::kTravErr] & 2) || n > (int64_t)a.cap_out)
        return fail(FGS_E_CAPACITY,
                    "tree_traverse: more than N + n_roots ids (duplicate roots or a cyclic tree)%s");
    if (!append && last != 0)
        return fail(FGS_E_INVALID, "tree_traverse: level_hint is below the tree's depth%s");
    *count = n;
    return FGS_OK;
}

int fgs_tree_traverse_result(int32_t N, int32_t C, int32_t n_roots, const void* ws, int64_t count,
                             int64_t* out, void* stream)
{
    if (N < 0 || C < 1 || n_roots < 0 || count < 0 || count > (int64_t)N + n_roots)
        return fail(FGS_E_INVALID, "tree_traverse_result: bad size%s");
    if (count == 0) return FGS_OK;
    if (!ws || !out) return fail(FGS_E_INVALID, "tree_traverse_result: null pointer%s");
    TravWs w = trav_carve(const_cast<void*>(ws), N, C, n_roots);
    return check(fgs::launch_traverse_emit(w.out, count, out, (hipStream_t)stream),
                 "tree_traverse_result");
}

int fgs_visible_flag(int32_t B, int32_t N, const float* xyz, const float* proj_matrix,
                     float padding, uint8_t* valid, float* depth, float* p_proj, void* stream)
{
    if (B < 0 || N < 0) return fail(FGS_E_INVALID, "visible_flag: negative size%s");
    if ((size_t)B * N > 0 && (!xyz || !proj_matrix || !valid || !depth || !p_proj))
        return fail(FGS_E_INVALID, "visible_flag: null pointer%s");
    return check(fgs::launch_visible_flag(B, N, xyz, proj_matrix, padding, valid, depth, p_proj,
                                          (hipStream_t)stream),
                 "visible_flag");
}

int fgs_gather_attributes(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                          const float* scaling, const float* rotation, const float* opacity,
                          const float* colors, const float* shs, int32_t sh_floats,
                          float* xyz_out, float* scaling_out, float* rotation_out,
                          float* opacity_out, float* colors_out, float* shs_out, int32_t* bad_index,
                          void* stream)
{
    if (N < 0 || n_src < 0) return fail(FGS_E_INVALID, "gather: negative size%s");
    if (N > 0 && (!index || !xyz || !scaling || !rotation || !opacity || !colors || !xyz_out ||
                  !scaling_out || !rotation_out || !opacity_out || !colors_out || !bad_index))
        return fail(FGS_E_INVALID, "gather: null pointer%s");
    return check(fgs::launch_gather(N, index, n_src, xyz, scaling, rotation, opacity, colors, shs,
                                     sh_floats, xyz_out, scaling_out, rotation_out, opacity_out,
                                     colors_out, shs_out, bad_index, (hipStream_t)stream),
                 "gather_attributes");
}

int fgs_activate_gather(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                        const float* log_scaling, const float* rotation, const float* opacity_logit,
                        const float* features_dc, const float* features_rest, int32_t rest_coeffs,
                        int32_t active_sh_degree, const float* campos, float* xyz_out,
                        float* scaling_out, float* rotation_out, float* opacity_out, float* colors_out,
                        int32_t* bad_index, void* stream)
{
    if (N < 0 || n_src < 0) return fail(FGS_E_INVALID, "activate: negative size%s");
    if (active_sh_degree < 0 || active_sh_degree > 4)
        return fail(FGS_E_INVALID, "activate: SH degree must be in [0, 4]%s");
    const int need = (active_sh_degree + 1) * (active_sh_degree + 1) - 1;
    if (active_sh_degree > 0 && (rest_coeffs < need || !features_rest || !campos))
        return fail(FGS_E_INVALID, "activate: SH rest coefficients / camera centre missing for the degree%s");
    if (N > 0 && (!xyz || !log_scaling || !rotation || !opacity_logit || !features_dc ||
                  !scaling_out || !rotation_out || !opacity_out || !colors_out || !bad_index))
        return fail(FGS_E_INVALID, "activate: null pointer%s");
    if (!index && N > n_src) return fail(FGS_E_INVALID, "activate: more rows than the source holds%s");
    fgs::ActivateArgs a{};
    a.N = N;
    a.n_src = n_src;
    a.index = index;
    a.xyz = xyz;
    a.log_scale = log_scaling;
    a.rot = rotation;
    a.opacity_logit = opacity_logit;
    a.dc = features_dc;
    a.rest = features_rest;
    a.rest_floats = rest_coeffs * 3;
    a.deg = active_sh_degree;
    a.campos = campos;
    a.xyz_o = xyz_out;
    a.scale_o = scaling_out;
    a.rot_o = rotation_out;
    a.opacity_o = opacity_out;
    a.colors_o = colors_out;
    a.bad_index = bad_index;
    return check(fgs::launch_activate_gather(a, (hipStream_t)stream), "activate_gather");
}

int fgs_activate_backward(int64_t N, const float* xyz, const float* log_scaling, const float* rotation,
                          const float* opacity_logit, int32_t rest_coeffs, int32_t active_sh_degree,
                          const float* campos, const float* dL_dscaling, const float* dL_drotation,
                          const float* dL_dopacity, const float* dL_dcolors, float* dL_dlog_scaling,
                          float* dL_drotation_raw, float* dL_dopacity_logit, float* dL_dfeatures_dc,
                          float* dL_dfeatures_rest, void* stream)
{
    if (N < 0 || rest_coeffs < 0) return fail(FGS_E_INVALID, "activate_backward: negative size%s");
    if (active_sh_degree < 0 || active_sh_degree > 4)
        return fail(FGS_E_INVALID, "activate_backward: SH degree must be in [0, 4]%s");
    if (N == 0) return FGS_OK;
    const int need = (active_sh_degree + 1) * (active_sh_degree + 1) - 1;
    if (dL_dfeatures_rest && active_sh_degree > 0 && (rest_coeffs < need || !xyz || !campos))
        return fail(FGS_E_INVALID, "activate_backward: SH rest coefficients / camera centre missing for the degree%s");
    if ((dL_dlog_scaling && !log_scaling) || (dL_drotation_raw && !rotation) ||
        (dL_dopacity_logit && !opacity_logit))
        return fail(FGS_E_INVALID, "activate_backward: null input for a requested gradient%s");
    fgs::ActivateBwdArgs a{};
    a.N = N;
    a.xyz = xyz;
    a.log_scale = log_scaling;
    a.rot = rotation;
    a.opacity_logit = opacity_logit;
    a.rest_floats = rest_coeffs * 3;
    a.deg = active_sh_degree;
    a.campos = campos;
    a.g_scale = dL_dscaling;
    a.g_rot = dL_drotation;
    a.g_opacity = dL_dopacity;
    a.g_colors = dL_dcolors;
    a.d_log_scale = dL_dlog_scaling;
    a.d_rot = dL_drotation_raw;
    a.d_opacity_logit = dL_dopacity_logit;
    a.d_dc = dL_dfeatures_dc;
    a.d_rest = dL_dfeatures_rest;
    return check(fgs::launch_activate_bwd(a, (hipStream_t)stream), "activate_backward");
}

/* ---- batched multi-view path ------------------------------------------- */

static int batch_validate(const fgs_batch_view* views, int32_t V, int32_t P)
{
    if (!views) return fail(FGS_E_INVALID, "null views%s");
    if (V <= 0 || V > fgs::kBatchMaxViews) return fail(FGS_E_INVALID, "views must be in [1, 64]%s");
    if (P < 0 || (uint64_t)P > fgs::kBatchIdxMask + 1)
        return fail(FGS_E_INVALID, "a batch holds at most 2^26 Gaussians%s");
    if ((uint64_t)V * (uint64_t)fgs::batch_p64(P) > 0x7FFFFFFFull)
        return fail(FGS_E_INVALID, "views * Gaussians (rounded up to 64) must stay below 2^31%s");
    for (int v = 0; v < V; ++v) {
        const fgs_batch_view& w = views[v];
        if (w.image_width <= 0 || w.image_height <= 0)
            return fail(FGS_E_INVALID, "image size must be positive%s");
        if ((w.image_width + 15) / 16 > 65535 || (w.image_height + 15) / 16 > 65535)
            return fail(FGS_E_INVALID, "image larger than 65535 tiles per side%s");
        if ((uint64_t)((w.image_width + 15) / 16) * (uint64_t)((w.image_height + 15) / 16) >=
            (uint64_t)fgs::kBatchMaxViewTiles)
            return fail(FGS_E_INVALID, "a batch view holds at most 2^26 tiles%s");
    }
    return FGS_OK;
}

int fgs_batch_sizes_query(const fgs_batch_view* views, int32_t V, int32_t P, fgs_batch_sizes* out)
{
    int rc = batch_validate(views, V, P);
    if (rc) return rc;
    if (!out) return fail(FGS_E_INVALID, "null output%s");
    int64_t T = 0, npix = 0, Tmax = 0;
    for (int v = 0; v < V; ++v) {
        const int gx = (views[v].image_width + 15) / 16, gy = (views[v].image_height + 15) / 16;
        T += (int64_t)gx * gy;
        Tmax = (int64_t)gx * gy > Tmax ? (int64_t)gx * gy : Tmax;
        npix += (int64_t)views[v].image_width * views[v].image_height;
    }
    if (T > 0xFFFFFFFELL) return fail(FGS_E_INVALID, "too many tiles in the batch%s");
    out->views = V;
    out->P = P;
    out->tiles = T;
    out->max_view_tiles = Tmax;
    out->pixels = npix;
    out->plan_bytes = align256(sizeof(fgs::BatchView) * (size_t)V);
    const size_t ne = (size_t)V * fgs::batch_p64(P);   // elements (batch_elem layout)
    carve_geom(nullptr, (int)ne, &out->geom_bytes, true);
    carve_image_n(nullptr, (size_t)npix, (size_t)T, &out->image_bytes);
    out->backward_ws_bytes = align256((ne > 0 ? ne : 1) * 3 * sizeof(float)) +
                             align256((ne > 0 ? ne : 1) * sizeof(float));
    return FGS_OK;
}

size_t fgs_batch_binning_bytes(int64_t K_cap)
{
    size_t t = 0;
    carve_binning(nullptr, K_cap, &t);
    return t;
}

int fgs_batch_plan(const fgs_batch_view* views, int32_t V, void* plan, void* stream)
{
    int rc = batch_validate(views, V, 0);
    if (rc) return rc;
    if (!plan) return fail(FGS_E_INVALID, "null plan buffer%s");
    std::vector<fgs::BatchView> bv((size_t)V);
    uint32_t tile_base = 0;
    uint64_t pix_base = 0;
    for (int v = 0; v < V; ++v) {
        const fgs_batch_view& w = views[v];
        fgs::BatchView& b = bv[(size_t)v];
        memcpy(b.view, w.viewmatrix, sizeof(b.view));
        memcpy(b.proj, w.projmatrix, sizeof(b.proj));
        memcpy(b.campos, w.campos, sizeof(b.campos));
        memcpy(b.bg, w.bg, sizeof(b.bg));
        b.tanx = w.tanfovx;
        b.tany = w.tanfovy;
        b.W = w.image_width;
        b.H = w.image_height;
        b.gx = (b.W + 15) / 16;
        b.gy = (b.H + 15) / 16;
        b.fy = (float)b.H / (2.0f * w.tanfovy);   // as make_pre_args
        b.fx = (float)b.W / (2.0f * w.tanfovx);
        b.tile_base = tile_base;
        b.pad = 0;
        b.pix_base = pix_base;
        tile_base += (uint32_t)(b.gx * b.gy);
        pix_base += (uint64_t)b.W * (uint64_t)b.H;
    }
    hipStream_t st = (hipStream_t)stream;
    if ((rc = check(hipMemcpyAsync(plan, bv.data(), sizeof(fgs::BatchView) * bv.size(),
                                   hipMemcpyHostToDevice, st), "plan upload")))
        return rc;
    return check(hipStreamSynchronize(st), "plan upload");
}

int fgs_batch_offsets(const fgs_batch_sizes* sz, int64_t K_cap, size_t* out4)
{
    if (!sz || !out4) return fail(FGS_E_INVALID, "null sizes or output%s");
    if (K_cap < 0) return fail(FGS_E_INVALID, "negative capacity%s");
    static char probe[256];
    Binning bn = carve_binning(probe, K_cap, nullptr);
    Image im = carve_image_n(probe, (size_t)sz->pixels, (size_t)sz->tiles, nullptr);
    out4[0] = (size_t)((char*)batch_pairs(bn, K_cap).keysA - probe);
    out4[1] = (size_t)((char*)im.ranges - probe);
    out4[2] = (size_t)((char*)im.final_T - probe);
    out4[3] = (size_t)((char*)im.n_contrib - probe);
    return FGS_OK;
}

static int batch_common(const fgs_raster_settings* s, const fgs_gaussians* g, const fgs_batch_sizes* sz)
{
    if (!s || !g || !sz) return fail(FGS_E_INVALID, "null settings, gaussians or sizes%s");
    if (g->P != sz->P) return fail(FGS_E_INVALID, "gaussians.P differs from the plan's P%s");
    if (g->P > 0 && (!g->means3D || !g->opacities))
        return fail(FGS_E_INVALID, "null means3D or opacities%s");
    if ((g->shs == nullptr) == (g->colors_precomp == nullptr) && g->P > 0)
        return fail(FGS_E_INVALID, "provide exactly one of shs / colors_precomp%s");
    if (g->P > 0 && g->shs && (g->M < (s->sh_degree + 1) * (s->sh_degree + 1) || s->sh_degree > 3 || s->sh_degree < 0))
        return fail(FGS_E_INVALID, "sh_degree must be in [0,3] with M >= (deg+1)^2%s");
    const bool sr = g->scales && g->rotations;
    if (g->P > 0 && (sr == (g->cov3D_precomp != nullptr)))
        return fail(FGS_E_INVALID, "provide exactly one of scales+rotations / cov3D_precomp%s");
    return FGS_OK;
}

static fgs::PreArgs batch_pre_args(const fgs_raster_settings* s, const fgs_gaussians* g)
{
    fgs_raster_settings c = *s;
    c.image_width = 16;   // per-view fields come from the plan
    c.image_height = 16;
    c.tanfovx = 1.f;
    c.tanfovy = 1.f;
    return make_pre_args(&c, g);
}

int fgs_batch_forward(const fgs_raster_settings* s, const fgs_gaussians* g, const fgs_batch_sizes* sz,
                      const void* plan, void* geom, void* image, void* binning, int64_t K_cap,
                      int32_t* radii, float* out_color, int32_t* out_point_id, float* out_weight_pixel,
                      float* out_point_weight, uint64_t* counts, void* stream)
{
    int rc = batch_common(s, g, sz);
    if (rc) return rc;
    if (!plan || !counts || !image || !out_color || !out_point_id || !out_weight_pixel)
        return fail(FGS_E_INVALID, "null plan or output%s");
    if (K_cap < 0) return fail(FGS_E_INVALID, "negative capacity%s");
    hipStream_t st = (hipStream_t)stream;
    const int V = sz->views, P = g->P;
    const size_t n = (size_t)V * (size_t)P;
    const size_t ne = (size_t)V * fgs::batch_p64(P);   // elements (batch_elem layout)
    const int T = (int)sz->tiles;
    const fgs::BatchView* views = static_cast<const fgs::BatchView*>(plan);
    if ((rc = check(fgs::launch_fill_u32(counts, 0u, 8, st), "counts clear"))) return rc;
    Image im = carve_image_n(image, (size_t)sz->pixels, (size_t)T, nullptr);
    Geom gm = carve_geom(geom, (int)(ne > 0 ? ne : 1), nullptr, true);
    fgs::BlendArgs b{};
    b.ranges = im.ranges;
    b.tile_vis = im.tile_vis;
    b.batch = g_shape[0];
    b.walk = g_walk;
    b.fwd_layout = fwd_layout_for((int64_t)T);
    b.xcd_chunk = g_xcd_chunk;
    b.id_stride = (uint32_t)P;
    const uint32_t* sorted_keys = nullptr;
    if (n > 0) {
        if (!geom || !radii || (K_cap > 0 && !binning) || !out_point_weight)
            return fail(FGS_E_INVALID, "null geometry / binning buffer or output%s");
        const fgs::PreArgs a = batch_pre_args(s, g);
        fgs::PreOut o{};
        o.radii = radii;
        o.means2D = gm.means2D;
        o.conic_opacity = gm.conic_opacity;
        o.rgb = gm.rgb;
        o.recs = gm.recs;             // depths and tile counts are not stored per element
        o.vkeys = gm.vkeysA;
        o.error_flag = (int*)(counts + 2);
        o.seg_cnt = gm.seg_cnt;
        o.vis_masks = gm.vis_masks;
        o.gid = gm.gid;
        o.grp_tiles = gm.grp_tiles;
        o.big_list = gm.big;
        o.big_count = gm.nbig;
        {
            ProfScope ps(ST_PREPROCESS, st);
            if ((rc = check(fgs::launch_fill_u32(gm.nbig, 0u, 2, st), "big count clear")) ||
                (rc = check(fgs::launch_preprocess_fwd_batch(a, views, V, o, st), "batch preprocess")))
                return rc;
            // each (view, 64-group) owns a run of backward gradient slots
            if ((rc = check(fgs::launch_scan_u32(gm.grp_tiles, ne / FGS_WAVE, gm.grp_base, gm.partials, st),
                            "batch slot bases")))
                return rc;
        }
        // (view, depth, index) order of the visible keys: bits 26..63, 5
        // passes -> result in B, their number in gm.nvis
        {
            ProfScope ps(ST_DEPTH_SORT, st);
            if ((rc = check(fgs::radix_sort_u64_segments(gm.vkeysA, gm.vkeysB, gm.seg_cnt, P, V,
                                                         fgs::kBatchIdxBits, 5, gm.hist, gm.partials,
                                                         gm.nvis, st), "batch depth sort")))
                return rc;
        }
        unsigned long long* cnt = (unsigned long long*)counts;
        {
            ProfScope ps(ST_OFFSETS, st);
            if ((rc = check(fgs::launch_sorted_records_batch(gm.vkeysB, gm.recs, n, gm.nvis, P, views, gm.tiles_s,
                                                             gm.rects_s, st), "batch records")))
                return rc;
            if ((rc = check(fgs::launch_scan_tiles_batch(gm.tiles_s, gm.offs, n, gm.nvis, gm.partials, cnt, st),
                            "batch offsets")))
                return rc;
        }
        if (K_cap > 0) {
            Binning bn = carve_binning(binning, K_cap, nullptr);
            const BatchPairs bp = batch_pairs(bn, K_cap);
            // keys carry the view's own tile number: the largest view sets the
            // passes, and the bits are split evenly over them (shorter digits
            // rank faster and leave longer runs to write); the first pass is
            // fused with the duplication and the result ends in A
            const int bits = bit_length((uint32_t)sz->max_view_tiles);
            const int passes = (bits + 7) / 8;
            const int db = (bits + passes - 1) / passes;
            const bool even = passes % 2 == 0;
            uint32_t* k1 = even ? bp.keysB : bp.keysA;
            uint32_t* v1 = even ? bp.valsB : bp.valsA;
            {
                ProfScope ps(ST_DUPLICATE, st);
                // the pair terms are free until the backward: chunk owners
                if ((rc = check(fgs::launch_duplicate_sort_batch(
                                    gm.vkeysB, gm.offs, gm.rects_s, gm.tiles_s, n, gm.nvis, cnt + 1, P, (size_t)K_cap,
                                    reinterpret_cast<uint32_t*>(bn.part), fgs::kBatchKeyViewBits, db, k1, v1,
                                    bn.hist, bn.partials, st),
                                "batch duplicate")))
                    return rc;
            }
            if (passes > 1) {
                ProfScope ps(ST_TILE_SORT, st);
                if ((rc = check(fgs::radix_sort_kv32(k1, v1, even ? bp.keysA : bp.keysB, even ? bp.valsA : bp.valsB,
                                                     (size_t)K_cap, fgs::kBatchKeyViewBits + db, passes - 1, db,
                                                     bn.hist, bn.partials, st, cnt + 1), "batch tile sort")))
                    return rc;
            }
            sorted_keys = bp.keysA;
            b.pair_ids = bp.valsA;
            b.pair_shift = 0;
            b.vis_ids = bp.keysB;   // free once sorted
            b.kv_strips = bn.strips;
        }
        b.means2D = gm.means2D;
        b.colors = gm.rgb;   // the preprocess copies colors_precomp to the elements
        b.conic_opacity = gm.conic_opacity;
        b.gid = gm.gid;
    }
    {
        ProfScope ps(ST_RANGES, st);
        if ((rc = check(fgs::launch_ranges_batch(sorted_keys, sorted_keys ? (size_t)K_cap : 0,
                                                 (const unsigned long long*)counts + 1, views, im.ranges, T, st),
                        "batch ranges")))
            return rc;
    }
    {
        ProfScope ps(ST_BLEND_FWD, st);
        if (g_queue_chunk > 0 && b.fwd_layout == 1 && T > 0) {
            if ((rc = check(fgs::build_tile_queue(im.ranges, nullptr, (uint32_t)T, (uint32_t)g_queue_chunk,
                                                  g_queue_by_cost, im.qkeysA, im.qkeysB, im.qhist, im.qpart,
                                                  im.qhead, st), "batch tile queue")))
                return rc;
            b.queue_order = im.qkeysB;
            b.queue_head = im.qhead;
            b.queue_chunk = (uint32_t)g_queue_chunk;
        }
        if ((rc = check(fgs::launch_blend_fwd_batch(b, views, V, T, im.final_T, im.n_contrib, im.vis_last,
                                                    out_color,
                                                    out_point_id, out_weight_pixel, out_point_weight,
                                                    st), "batch blend")))
            return rc;
    }
    return sync_check(s, st, "batch forward");
}

int fgs_batch_backward(const fgs_raster_settings* s, const fgs_gaussians* g, const fgs_batch_sizes* sz,
                       const void* plan, const int32_t* radii, const void* geom, const void* image,
                       const void* binning, int64_t K_cap, const float* dL_dpix, float* dL_dmeans2D,
                       float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D,
                       float* dL_dsh, float* dL_dscales, float* dL_drot, void* ws, void* stream)
{
    int rc = batch_common(s, g, sz);
    if (rc) return rc;
    const int V = sz->views, P = g->P;
    const size_t n = (size_t)V * fgs::batch_p64(P);   // elements (batch_elem layout)
    if (P == 0) return FGS_OK;
    if (!plan || !radii || !geom || !image || (K_cap > 0 && !binning) || !dL_dpix || !dL_dmeans2D || !dL_dcolors ||
        !dL_dopacity || !dL_dmeans3D || !dL_dcov3D || !ws)
        return fail(FGS_E_INVALID, "null input, output or workspace%s");
    if (g->shs && !dL_dsh) return fail(FGS_E_INVALID, "null dL_dsh%s");
    if (!g->cov3D_precomp && (!dL_dscales || !dL_drot)) return fail(FGS_E_INVALID, "null dL_dscales / dL_drot%s");
    hipStream_t st = (hipStream_t)stream;
    const int T = (int)sz->tiles;
    const fgs::BatchView* views = static_cast<const fgs::BatchView*>(plan);
    Geom gm = carve_geom(const_cast<void*>(geom), (int)n, nullptr, true);
    Image im = carve_image_n(const_cast<void*>(image), (size_t)sz->pixels, (size_t)T, nullptr);
    // per-element blend gradients: the gather stores those of every visible
    // element (nothing reads the others); without pairs they are cleared
    float* dconic = static_cast<float*>(ws);
    float* dop_view = reinterpret_cast<float*>(static_cast<char*>(ws) + align256(n * 3 * sizeof(float)));
    if (K_cap == 0 && (rc = check(fgs::launch_clear_elements(gm.vis_masks, n / FGS_WAVE, dL_dmeans2D, dconic,
                                                             dL_dcolors, dop_view, st), "element clear")))
        return rc;
    if (K_cap > 0) {
        Binning bn = carve_binning(const_cast<void*>(binning), K_cap, nullptr);
        const BatchPairs bp = batch_pairs(bn, K_cap);
        fgs::BlendArgs b{};
        b.ranges = im.ranges;
        b.tile_vis = im.tile_vis;
        b.pair_ids = bp.valsA;
        b.pair_shift = 0;
        b.vis_ids = bp.keysB;
        b.kv_strips = bn.strips;
        b.recs = gm.recs;
        b.slot_base = gm.grp_base;
        b.part = bn.part;
        b.mark = bn.mark;
        b.part_cap = (size_t)K_cap;
        b.means2D = gm.means2D;
        b.colors = gm.rgb;
        b.conic_opacity = gm.conic_opacity;
        b.bwd_batch = g_shape[1];
        b.bwd_layout = g_bwd_layout;
        b.xcd_chunk = g_xcd_chunk;
        ProfScope ps(ST_BLEND_BWD, st);
        if (g_queue_chunk > 0 && b.bwd_layout == 1 && T > 0) {
            if ((rc = check(fgs::build_tile_queue(im.ranges, im.tile_vis, (uint32_t)T, (uint32_t)g_queue_chunk,
                                                      g_queue_by_cost, im.qkeysA, im.qkeysB, im.qhist,
                                                      im.qpart, im.qhead, st), "batch tile queue")))
                return rc;
            b.queue_order = im.qkeysB;
            b.queue_head = im.qhead;
            b.queue_chunk = (uint32_t)g_queue_chunk;
        }
        // the visible elements in element order (the depth order's record
        // arrays are free after the forward)
        if ((rc = check(fgs::launch_visible_list(gm.vis_masks, n / FGS_WAVE, gm.offs, gm.tiles_s, gm.partials, st),
                        "visible elements")))
            return rc;
        const fgs::GradGather gg{ nullptr, nullptr, nullptr, (size_t)V * (size_t)P, gm.nvis, gm.tiles_s, gm.recs,
                                  gm.grp_base, gm.big, gm.nbig };
        if ((rc = check(fgs::launch_blend_bwd_batch(b, gg, views, V, T, im.final_T, im.vis_last, dL_dpix,
                                                    dL_dmeans2D, dconic, dop_view, dL_dcolors,
                                                    s->grad_mode == 1, st), "batch blend backward")))
            return rc;
    }
    const fgs::PreArgs a = batch_pre_args(s, g);
    fgs::PreBwdIn in{ radii, dL_dmeans2D, dconic, dL_dcolors, gm.vis_masks, gm.rgb };
    fgs::PreBwdOut o{};
    o.dL_dopacity = dL_dopacity;
    o.dL_dmeans3D = dL_dmeans3D;
    o.dL_dcov3D = dL_dcov3D;
    o.dL_dsh = g->shs ? dL_dsh : nullptr;
    o.dL_dscales = g->cov3D_precomp ? nullptr : dL_dscales;
    o.dL_drot = g->cov3D_precomp ? nullptr : dL_drot;
    {
        ProfScope ps(ST_PREPROCESS_BWD, st);
        if ((rc = check(fgs::launch_preprocess_bwd_batch(a, views, V, in, dop_view, o, st),
                        "batch preprocess backward")))
            return rc;
    }
    return sync_check(s, st, "batch backward");
}

}  // extern "C"

namespace fgs {
void prof_kernel(bool backward, bool begin, hipStream_t s)
{
    const int k = backward ? ST_BLEND_BWD_KERNEL : ST_BLEND_FWD_KERNEL;
    if (begin) prof_begin(k, s);
    else prof_end(k, s);
}
}  // namespace fgs
