// Local time stepping, host side: the C-ABI of the level stepper's operations (the methods of
// wadi2d/solver/gpu/levels.py GpuLevels, which bind levels.cu and the tile-list kernels). Included at the end of
// step.hip. The kernels are in levels_kernels.hpp.

namespace {

void lv_drop(w2d_handle* H, void* p) {
    if (!p) return;
    H->owned.erase(std::remove(H->owned.begin(), H->owned.end(), p), H->owned.end());
    (void)hipFree(p);
}

// The list `id` as the kernels take it, or an error for an id this handle has not given out or a list of the other
// kind (tiles against edge rows).
int lv_list(w2d_handle* H, int32_t id, int width, const int** out, int* n) {
    if (!H->lv_r) return fail(1, "level: w2d_level_setup first");
    if (id < 0 || id >= (int)H->lv_lists.size() || !H->lv_lists[id]) return fail(1, "level: no such list");
    if (H->lv_width[id] != width)
        return fail(1, width == 1 ? "level: this operation takes a list of tiles" : "level: this operation takes edge rows");
    *out = H->lv_lists[id];
    *n = H->lv_count[id];
    return 0;
}

template <class R>
LevelArgs<R> level_args(w2d_handle* H, const int* list, int n, double dt, int flag) {
    LevelArgs<R> L{};
    L.A = step_args<R>(H, dt, 0);
    L.A.fresh = 1;
    L.list = list;
    L.n = n;
    L.dt = (R)dt;  // float32 kernels take np.float32(dt), as the uniform step's
    L.flag = flag;
    L.level = H->lv_level;
    L.h0 = (const R*)H->lv_h0;
    L.d0 = (const R*)H->lv_d0;
    L.sub = H->lv_sub;
    L.r = (R*)H->lv_r;
    L.est = (R*)H->lv_est;
    return L;
}

// One launch of an operation over list `id`: grid of one block per tile or edge row.
enum LevelOp { LV_FACES, LV_STRIPS, LV_LIMIT, LV_SAVE, LV_UPDATE, LV_COMMIT, LV_COMMIT_STRIPS, LV_ESTIMATE, LV_FLUX };

template <class R, bool FAST>
int level_launch(w2d_handle* H, LevelOp op, const int* list, int n, double dt, int flag) {
    if (n == 0) return 0;
    const LevelArgs<R> L = level_args<R>(H, list, n, dt, flag);
    switch (op) {
    case LV_FACES: faces_level_kernel<R, FAST><<<n, NT, 0, H->stream>>>(L); break;
    case LV_STRIPS: strips_kernel<R, FAST><<<n, NT, 0, H->stream>>>(L); break;
    case LV_LIMIT: limit_level_kernel<R, FAST><<<n, NT, 0, H->stream>>>(L); break;
    case LV_SAVE: save_level_kernel<R><<<n, NT, 0, H->stream>>>(L, (R*)H->lv_h0, (R*)H->lv_d0); break;
    case LV_UPDATE: update_level_kernel<R, FAST><<<n, NT, 0, H->stream>>>(L); break;
    case LV_COMMIT: commit_level_kernel<R><<<n, NT, 0, H->stream>>>(L); break;
    case LV_COMMIT_STRIPS: commit_strips_kernel<R><<<n, TILE, 0, H->stream>>>(L); break;
    case LV_ESTIMATE: estimate_kernel<R><<<n, TILE, 0, H->stream>>>(L); break;
    case LV_FLUX: flux_kernel<R, FAST><<<n, TILE, 0, H->stream>>>(L); break;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int level_op(w2d_handle* H, LevelOp op, int32_t id, double dt, int flag) {
    if (!H) return fail(1, "level: no handle");
    const bool rows = op == LV_STRIPS || op == LV_COMMIT_STRIPS || op == LV_ESTIMATE || op == LV_FLUX;
    const int* list;
    int n, s;
    if ((s = lv_list(H, id, rows ? 4 : 1, &list, &n))) return s;
    if (H->prec == 8) return level_launch<double, false>(H, op, list, n, dt, flag);
    if (H->fast) return level_launch<float, true>(H, op, list, n, dt, flag);
    return level_launch<float, false>(H, op, list, n, dt, flag);
}

}  // namespace

extern "C" {

int w2d_level_setup(w2d_handle* H) {
    if (!H) return fail(1, "level_setup: no handle");
    if (H->closure_n) return fail(1, "level_setup: the learned closure runs with the uniform step only");
    int s;
    if ((s = run_align(H))) return s;
    if ((s = materialize(H))) return s;  // the level kernels read the stored face momentum
    if (!H->lv_r) {
        const size_t cells = H->ncell * (size_t)H->prec;
        if ((s = dalloc(H, &H->lv_r, cells)) || (s = dalloc(H, &H->lv_h0, cells)) || (s = dalloc(H, &H->lv_d0, cells)) ||
            (s = dalloc(H, &H->lv_est, (size_t)4 * H->ntiles * TILE * H->prec)) ||
            (s = dalloc(H, (void**)&H->lv_level, (size_t)H->ntiles * sizeof(int))))
            return s;
    }
    HIP_TRY(hipStreamSynchronize(H->stream));
    return 0;
}

int w2d_level_set(w2d_handle* H, const int32_t* level) {
    if (!H || !level) return fail(1, "level_set: bad arguments");
    if (!H->lv_r) return fail(1, "level_set: w2d_level_setup first");
    for (int t = 0; t < H->ntiles; ++t)
        if (level[t] < 0 || level[t] > 30) return fail(1, "level_set: a level is out of range");
    HIP_TRY(hipMemcpy(H->lv_level, level, H->ntiles * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

int w2d_level_list(w2d_handle* H, const int32_t* a, int32_t n, int32_t width, int32_t* id) {
    if (!H || n < 0 || (n && !a) || !id || (width != 1 && width != 4)) return fail(1, "level_list: bad arguments");
    if (!H->lv_r) return fail(1, "level_list: w2d_level_setup first");
    for (int k = 0; k < n; ++k) {
        const int32_t* e = a + (size_t)k * width;
        if (width == 1) {
            if (e[0] <= 0 || e[0] >= H->ntiles) return fail(1, "level_list: tile out of range");
        } else if (e[0] <= 0 || e[0] >= H->ntiles || e[2] <= 0 || e[2] >= H->ntiles || e[1] < 0 || e[1] > 3 ||
                   e[3] < 0 || e[3] >= 4 * H->ntiles) {
            return fail(1, "level_list: an edge row is out of range");
        }
    }
    int* dev = nullptr;
    int s;
    if ((s = upload(H, &dev, a ? a : id, std::max<size_t>((size_t)n * width, 1)))) return s;
    H->lv_lists.push_back(dev);
    H->lv_count.push_back(n);
    H->lv_width.push_back(width);
    *id = (int32_t)H->lv_lists.size() - 1;
    return 0;
}

int w2d_level_lists_clear(w2d_handle* H) {
    if (!H) return fail(1, "level_lists_clear: no handle");
    HIP_TRY(hipStreamSynchronize(H->stream));
    for (int* p : H->lv_lists) lv_drop(H, p);
    H->lv_lists.clear();
    H->lv_count.clear();
    H->lv_width.clear();
    return 0;
}

int w2d_level_halo_state(w2d_handle* H, int32_t sub) {
    if (!H || sub < 0) return fail(1, "level_halo_state: bad arguments");
    H->lv_sub = sub;
    return 0;
}

int w2d_level_faces(w2d_handle* H, int32_t id, double dt, int32_t want_wave) {
    return level_op(H, LV_FACES, id, dt, want_wave != 0);
}
int w2d_level_strips(w2d_handle* H, int32_t id, double dt, int32_t want_wave) {
    return level_op(H, LV_STRIPS, id, dt, want_wave != 0);
}
int w2d_level_limit(w2d_handle* H, int32_t id, double dt) { return level_op(H, LV_LIMIT, id, dt, 0); }
int w2d_level_save(w2d_handle* H, int32_t id) { return level_op(H, LV_SAVE, id, 0.0, 0); }
int w2d_level_update(w2d_handle* H, int32_t id, double dt, int32_t want_plan) {
    return level_op(H, LV_UPDATE, id, dt, want_plan != 0);
}
int w2d_level_commit(w2d_handle* H, int32_t id) { return level_op(H, LV_COMMIT, id, 0.0, 0); }
int w2d_level_commit_strips(w2d_handle* H, int32_t id) { return level_op(H, LV_COMMIT_STRIPS, id, 0.0, 0); }
int w2d_level_estimate(w2d_handle* H, int32_t id) { return level_op(H, LV_ESTIMATE, id, 0.0, 0); }
int w2d_level_flux(w2d_handle* H, int32_t id, double dt) { return level_op(H, LV_FLUX, id, dt, 0); }

int w2d_level_clear(w2d_handle* H, int32_t what) {
    if (!H || what < 0 || what > 1) return fail(1, "level_clear: bad arguments");
    if (what == 0) {
        if (H->prec == 8)
            fill_kernel<double><<<64, 256, 0, H->stream>>>((double*)H->wave, (size_t)H->ntiles, 0.0);
        else
            fill_kernel<float><<<64, 256, 0, H->stream>>>((float*)H->wave, (size_t)H->ntiles, 0.0f);
    } else {
        fill_kernel<int><<<64, 256, 0, H->stream>>>(H->wet, (size_t)H->ntiles, 0);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int w2d_level_boundary(w2d_handle* H) {
    if (!H) return fail(1, "level_boundary: no handle");
    if (!H->nb_cells) return 0;
    const Cur<char> c = cur(H);  // the level run updates the current depth in place
    if (H->prec == 8)
        boundary_kernel<double><<<(H->nb_cells + 255) / 256, 256, 0, H->stream>>>(
            H->nb_cells, H->b_cell, H->b_ref, H->b_kind, H->b_d0, H->b_h0, (double*)c.h, (double*)c.d);
    else
        boundary_kernel<float><<<(H->nb_cells + 255) / 256, 256, 0, H->stream>>>(
            H->nb_cells, H->b_cell, H->b_ref, H->b_kind, H->b_d0, H->b_h0, (float*)c.h, (float*)c.d);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
