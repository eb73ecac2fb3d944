// batch.cpp -- uneven batches: slices of different geometries in one persistent launch.
//
// The reference handles one geometry per op call, held in process-global state
// (FP/transform.cu:15-18,20-62), and loops its slices on the host with one kernel launch
// per view (FP/transform.cu:233-255).  A batch here lists one plan per slice; its work
// tables (per-slice descriptors pointing at each plan's device tables, and prefix sums of
// the work items per slice) are built once, so every execute is a single kernel launch
// over all slices that allocates nothing.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <new>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kMaxWgsPerCu = 32;

void free_batch_device(BatchDevice &d)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    if (d.device >= 0) (void)hipSetDevice(d.device);
    (void)hipFree(d.fp_desc);
    (void)hipFree(d.fpa_desc);
    for (int k = 0; k < kFpNumCfg; ++k) (void)hipFree(d.fp_prefix[k]);
    for (int k = 0; k < kFpNumCfg; ++k) (void)hipFree(d.fp_order[k]);
    for (int k = 0; k < kAdjNumTiles; ++k) (void)hipFree(d.fpa_prefix[k]);
    if (cur >= 0) (void)hipSetDevice(cur);
    d = BatchDevice{};
}

fbct_status batch_device(fbct_batch *b, const BatchDevice **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(b->mu);
    for (const BatchDevice &d : b->dev)
        if (d.device == dev) {
            *out = &d;
            return FBCT_OK;
        }
    const size_t n = b->plans.size();
    std::vector<FpSliceDesc> fpd(n);
    std::vector<FpaSliceDesc> fpad(n);
    for (size_t s = 0; s < n; ++s) {
        fbct_plan *p = b->plans[s];
        const DeviceTables *t = nullptr;
        fbct_status st = device_tables(p, &t);
        if (st != FBCT_OK) return st;
        FpSliceDesc &f = fpd[s];
        f.ray_geo = t->ray_geo;
        f.ray_n = t->ray_n;
        for (int k = 0; k < kFpNumCfg; ++k) {
            f.major[k] = t->fp_major[k];
            f.slabs[k] = t->fp_slabs[k];
            f.rows[k] = t->fp_rows[k];
            f.nb[k] = p->d.fpb[k].nb;
            f.cap[k] = p->d.fpb[k].cap;
            f.nslab[k] = p->d.fpb[k].nslab;
        }
        f.img_off = b->img_off[s];
        f.sino_off = b->sino_off[s];
        f.nX = p->g.nX;
        f.nY = p->g.nY;
        f.nViews = p->g.nViews;
        f.nCols = p->g.nCols;
        f.scale = p->d.fp_scale;
        FpaSliceDesc &a = fpad[s];
        a.ray_geo = t->ray_geo;
        a.ray_n = t->ray_n;
        a.ray_inv = t->ray_inv;
        a.cs = t->view_cs;
        a.img_off = b->img_off[s];
        a.sino_off = b->sino_off[s];
        a.nX = p->g.nX;
        a.nY = p->g.nY;
        a.nViews = p->g.nViews;
        a.nCols = p->g.nCols;
        a.scale = p->d.fp_scale;
        a.r = p->d.fpa_r;
        a.dx = p->g.dx;
        a.R = p->g.R;
        a.c_scale = p->d.fpa_c_scale;
        a.c_off = p->d.fpa_c_off;
        a.c_abs = std::fabs(p->d.fpa_c_scale);
        a.far_x = p->d.fpa_far_x;
        a.far_y = p->d.fpa_far_y;
    }
    BatchDevice d;
    d.device = dev;
    e = hipDeviceGetAttribute(&d.num_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || d.num_cu <= 0) return fail(FBCT_ERR_DEVICE, "compute unit count: %s", hipGetErrorString(e));
    fbct_status st = upload_vec(&d.fp_desc, fpd);
    if (st == FBCT_OK) st = upload_vec(&d.fpa_desc, fpad);
    for (int k = 0; k < kFpNumCfg && st == FBCT_OK; ++k) st = upload_vec(&d.fp_prefix[k], b->fp_prefix[k]);
    for (int k = 0; k < kFpNumCfg && st == FBCT_OK; ++k) st = upload_vec(&d.fp_order[k], b->fp_order[k]);
    for (int k = 0; k < kAdjNumTiles && st == FBCT_OK; ++k) st = upload_vec(&d.fpa_prefix[k], b->fpa_prefix[k]);
    if (st != FBCT_OK) {
        free_batch_device(d);
        return st;
    }
    if (b->dev.capacity() == 0) b->dev.reserve(16);
    if (b->dev.size() == b->dev.capacity()) {
        free_batch_device(d);
        return fail(FBCT_ERR_UNSUPPORTED, "batch used on more than %zu devices", b->dev.capacity());
    }
    b->dev.push_back(d);
    *out = &b->dev.back();
    return FBCT_OK;
}

fbct_status check_shape(const fbct_batch_shape &s)
{
    if (s.fp_rays != 64 && s.fp_rays != 32)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "batch shape: fp_rays must be 64 or 32 (got %d)", s.fp_rays);
    if (s.fp_wgs_per_cu < 0 || s.fp_wgs_per_cu > kMaxWgsPerCu || s.adj_wgs_per_cu < 0 ||
        s.adj_wgs_per_cu > kMaxWgsPerCu)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "batch shape: workgroups per CU must be in 0..%d", kMaxWgsPerCu);
    for (int k = 0; k < kAdjNumTiles; ++k)
        if (kAdjTileX[k] == s.adj_tile_x) {
            if (s.adj_stage_views != 0 && s.adj_stage_views != 4 && s.adj_stage_views != 8)
                return fail(FBCT_ERR_INVALID_ARGUMENT, "batch shape: adj_stage_views must be 0, 4 or 8 (got %d)",
                            s.adj_stage_views);
            return FBCT_OK;
        }
    return fail(FBCT_ERR_INVALID_ARGUMENT, "batch shape: adj_tile_x must be 8, 16, 32 or 64 (got %d)", s.adj_tile_x);
}

fbct_status check_exec(fbct_batch *b, const void *in, int ti, const void *out, int to, const char *what)
{
    if (!b) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null batch", what);
    if (!in || !out) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null data pointer", what);
    if ((ti != FBCT_F32 && ti != FBCT_BF16) || (to != FBCT_F32 && to != FBCT_BF16))
        return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: unknown storage type", what);
    return FBCT_OK;
}

}  // namespace
}  // namespace fbct

extern "C" {

fbct_status fbct_batch_create(fbct_plan *const *plans, int n_slices, fbct_batch **out)
{
    using fbct::fail;
    if (!out) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_create: null output pointer");
    *out = nullptr;
    if (!plans) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_create: null plan list");
    if (n_slices < 1) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_create: need at least one slice");
    for (int s = 0; s < n_slices; ++s)
        if (!plans[s]) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_create: plan %d is null", s);
    fbct_batch *b = new (std::nothrow) fbct_batch();
    if (!b) return fail(FBCT_ERR_OUT_OF_MEMORY, "batch allocation");
    try {
        b->plans.assign(plans, plans + n_slices);
        b->img_off.assign(1, 0);
        b->sino_off.assign(1, 0);
        for (int k = 0; k < fbct::kFpNumCfg; ++k) b->fp_prefix[k].assign(1, 0);
        for (int k = 0; k < fbct::kAdjNumTiles; ++k) b->fpa_prefix[k].assign(1, 0);
        int64_t fp_items[fbct::kFpNumCfg] = {}, fpa_items[fbct::kAdjNumTiles] = {};
        for (int s = 0; s < n_slices; ++s) {
            const fbct_plan &p = *plans[s];
            b->img_off.push_back(b->img_off.back() + (int64_t)p.g.nX * p.g.nY);
            b->sino_off.push_back(b->sino_off.back() + (int64_t)p.g.nViews * p.g.nCols);
            for (int k = 0; k < fbct::kFpNumCfg; ++k) {
                fp_items[k] += (int64_t)p.g.nViews * p.d.fpb[k].nb;
                b->fp_cap_max[k] = std::max(b->fp_cap_max[k], p.d.fpb[k].cap);
            }
            const bool cells = fbct::adj_cells_fit(p.d.fpa_far_x, p.d.fpa_far_y);
            for (int k = 0; k < fbct::kAdjNumTiles; ++k) {
                const int tx = fbct::kAdjTileX[k], ty = 256 / tx;
                fpa_items[k] += cells ? (int64_t)fbct::adj_cell_tiles(p.g.nX) * fbct::adj_cell_tiles(p.g.nY)
                                      : (int64_t)((p.g.nX + tx - 1) / tx) * ((p.g.nY + ty - 1) / ty);
            }
            for (int64_t v : fp_items)
                if (v > INT_MAX) {
                    delete b;
                    return fail(FBCT_ERR_UNSUPPORTED, "fbct_batch_create: more than 2^31 FP work items");
                }
            for (int64_t v : fpa_items)
                if (v > INT_MAX) {
                    delete b;
                    return fail(FBCT_ERR_UNSUPPORTED, "fbct_batch_create: more than 2^31 adjoint work items");
                }
            for (int k = 0; k < fbct::kFpNumCfg; ++k) b->fp_prefix[k].push_back((int)fp_items[k]);
            for (int k = 0; k < fbct::kAdjNumTiles; ++k) b->fpa_prefix[k].push_back((int)fpa_items[k]);
            b->fp_samples += p.fp_samples;
            b->updates += (int64_t)p.g.nX * p.g.nY * p.g.nViews;
        }
        // FP items of all slices: those whose largest tile exceeds the rows of
        // kFpWgsPerCu workgroups per CU first, then the rest, each in item order
        for (int k = 0; k < fbct::kFpNumCfg; ++k) {
            const int total = b->fp_prefix[k][n_slices];
            const size_t row_bytes = static_cast<size_t>(fbct::fp_pitch(k)) * sizeof(float);
            const int small_cap = static_cast<int>(std::min<size_t>(
                static_cast<size_t>(b->fp_cap_max[k]), (160 * 1024 / fbct::kFpWgsPerCu) / row_bytes));
            std::vector<uint8_t> big(static_cast<size_t>(total), 0);
            for (int s = 0; s < n_slices; ++s) {
                const fbct_plan &p = *plans[s];
                const int base = b->fp_prefix[k][s];
                for (size_t q = 0; q < p.fp_order[k].size(); ++q)
                    if (p.fp_order_rows[k][q] > small_cap) big[static_cast<size_t>(base + p.fp_order[k][q])] = 1;
            }
            std::vector<int32_t> &order = b->fp_order[k];
            order.clear();
            order.reserve(static_cast<size_t>(total));
            for (int g = 0; g < total; ++g)
                if (big[g]) order.push_back(g);
            b->fp_split[k] = static_cast<int>(order.size());
            for (int g = 0; g < total; ++g)
                if (!big[g]) order.push_back(g);
            b->fp_small_cap[k] = small_cap;
        }
        b->shape = fbct_batch_shape{64, 0, 16, 0, 0};
    } catch (const std::bad_alloc &) {
        delete b;
        return fail(FBCT_ERR_OUT_OF_MEMORY, "batch tables");
    }
    *out = b;
    return FBCT_OK;
}

fbct_status fbct_batch_destroy(fbct_batch *batch)
{
    if (!batch) return FBCT_OK;
    for (fbct::BatchDevice &d : batch->dev) fbct::free_batch_device(d);
    delete batch;
    return FBCT_OK;
}

int fbct_batch_size(const fbct_batch *batch)
{
    return batch ? (int)batch->plans.size() : -1;
}

fbct_status fbct_batch_layout(const fbct_batch *batch, int64_t *image_offsets, int64_t *sino_offsets)
{
    if (!batch) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_layout: null batch");
    if (image_offsets) std::copy(batch->img_off.begin(), batch->img_off.end(), image_offsets);
    if (sino_offsets) std::copy(batch->sino_off.begin(), batch->sino_off.end(), sino_offsets);
    return FBCT_OK;
}

fbct_status fbct_batch_work(const fbct_batch *batch, int64_t *fp_samples, int64_t *updates)
{
    if (!batch) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_work: null batch");
    if (fp_samples) *fp_samples = batch->fp_samples;
    if (updates) *updates = batch->updates;
    return FBCT_OK;
}

fbct_status fbct_batch_get_shape(const fbct_batch *batch, fbct_batch_shape *shape)
{
    if (!batch || !shape) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_get_shape: null pointer");
    *shape = batch->shape;
    return FBCT_OK;
}

fbct_status fbct_batch_set_shape(fbct_batch *batch, const fbct_batch_shape *shape)
{
    if (!batch || !shape) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_set_shape: null pointer");
    fbct_status s = fbct::check_shape(*shape);
    if (s != FBCT_OK) return s;
    batch->shape = *shape;
    return FBCT_OK;
}

fbct_status fbct_batch_upload(fbct_batch *batch)
{
    if (!batch) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_upload: null batch");
    const fbct::BatchDevice *d = nullptr;
    return fbct::batch_device(batch, &d);
}

fbct_status fbct_batch_fp(fbct_batch *batch, const void *images, int image_dtype, void *sinos, int sino_dtype,
                          void *stream)
{
    fbct_status s = fbct::check_exec(batch, images, image_dtype, sinos, sino_dtype, "fbct_batch_fp");
    if (s != FBCT_OK) return s;
    const fbct::BatchDevice *d = nullptr;
    if ((s = fbct::batch_device(batch, &d)) != FBCT_OK) return s;
    return fbct::launch_fp_batch(*batch, *d, images, static_cast<fbct_dtype>(image_dtype), sinos,
                                 static_cast<fbct_dtype>(sino_dtype), static_cast<hipStream_t>(stream));
}

fbct_status fbct_batch_fp_adjoint(fbct_batch *batch, const void *sinos, int sino_dtype, void *images,
                                  int image_dtype, void *stream)
{
    fbct_status s = fbct::check_exec(batch, sinos, sino_dtype, images, image_dtype, "fbct_batch_fp_adjoint");
    if (s != FBCT_OK) return s;
    const fbct::BatchDevice *d = nullptr;
    if ((s = fbct::batch_device(batch, &d)) != FBCT_OK) return s;
    return fbct::launch_fp_adjoint_batch(*batch, *d, sinos, static_cast<fbct_dtype>(sino_dtype), images,
                                         static_cast<fbct_dtype>(image_dtype), static_cast<hipStream_t>(stream));
}

}  // extern "C"
