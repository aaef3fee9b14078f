// mz_step.hip -- multi-step drivers: the 1-D Couette step (ntc! in every
// cell, convection, sort, compute_props_sorted!; test/test_1D_couette.jl:66-97)
// launched kernel by kernel or replayed as a captured hipGraph.
//
// A captured graph holds two steps: the sort writes into the second particle
// buffer, so two steps bring the buffers back to where the capture started.
// The random tags of captured kernels come from a device step counter that the
// graph's last node advances, so every replay draws fresh numbers in the same
// (step << 4) | phase layout as eager launches.
#include "mz_internal.h"

#include <cstdio>
#include <cstring>
#include <vector>

__global__ void k_set_step(uint64_t *step, uint64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}
__global__ void k_step_inc(uint64_t *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

void mz_graph_release(mz_store *h) {
    if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
    if (h->graph) hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_key.clear();
    if (h->d_step) hipFree(h->d_step);
    h->d_step = nullptr;
}

// one step on the current buffers; the host-side buffer swap happens here
// With avg_inv > 0 the step is one of the averaging phase of
// simulations/1D/couette_with_surface_quantities.jl:73-115: convection also
// computes the wall properties and adds them to the time average
// (avg_props! of SurfProps), and the cell properties are computed every step
// and added to theirs (avg_props! of PhysProps), each with weight avg_inv.
static int enqueue_step(mz_store *h, const double *dV, double dt, int do_props, uint64_t step,
                        const uint64_t *step_ptr, double avg_inv = 0.0, double mass = 0.0) {
    int r = mz_launch_ntc_equal(h, 1, 1, dt, dV, 1e-16, 1, (step << 4) | 1u, 0, 1, h->n_cells, step_ptr);
    if (r) return r;
    if (avg_inv > 0.0) {
        r = mz_launch_convect_surf(h, 1, dt, (step << 4) | 2u, 0, true, 0, mass, step_ptr, avg_inv);
    } else {
        r = mz_launch_convect(h, 1, dt, (step << 4) | 2u, 0, true, 0, step_ptr);
    }
    if (r) return r;
    r = mz_launch_sort(h, 1, 1);
    if (r) return r;
    mz_swap_buffers(h->sp[0]);
    if (do_props || avg_inv > 0.0) {
        r = mz_launch_props(h, 1, 0, avg_inv);
        if (r) return r;
    }
    return 0;
}

extern "C" int mz_couette_last_path(mz_store *h) { return h->last_path; }

static int couette_run(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                       int32_t equal_weight, int32_t props_every, const double *masses, int32_t mode,
                       double avg_inv);

extern "C" int mz_couette_steps(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                                int32_t equal_weight, int32_t props_every, const double *masses, int32_t mode) {
    return couette_run(h, n_steps, step0, dt, V, equal_weight, props_every, masses, mode, 0.0);
}

extern "C" int mz_couette_avg_steps(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                                    const double *masses, int32_t mode, double n_avg_timesteps) {
    MZ_CHECK(n_avg_timesteps > 0.0, "couette_avg_steps: n_avg_timesteps must be positive");
    MZ_CHECK(mode == 0 || mode == 1 || mode == 2, "couette_avg_steps: mode is 0 (eager), 1 (graph) or 2 (slab)");
    return couette_run(h, n_steps, step0, dt, V, 1, 1, masses, mode, 1.0 / n_avg_timesteps);
}

static int couette_run(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                       int32_t equal_weight, int32_t props_every, const double *masses, int32_t mode,
                       double avg_inv) {
    MZ_CHECK(h->n_species == 1, "couette_steps: single-species stores only");
    MZ_CHECK(h->has_grid && h->has_walls, "couette_steps: grid or walls not set");
    MZ_CHECK(h->inter[0].set, "couette_steps: interaction not set");
    MZ_CHECK(masses != nullptr, "couette_steps: masses is NULL");
    MZ_CHECK(n_steps >= 0, "couette_steps: negative step count");
    // uploads only of what changed since the last call (the device copies
    // are written by nothing else than these uploads)
    if (!h->masses_valid || h->masses_up[0] != masses[0]) {
        MZ_HIP(hipMemcpy(h->d_masses, masses, sizeof(double), hipMemcpyHostToDevice));
        h->masses_up[0] = masses[0];
        h->masses_valid = true;
    }
    const double *dV = h->d_V;
    if (V) {
        double *v = (double *)mz_scratch(h, "step_V", h->n_cells * sizeof(double));
        MZ_CHECK(v != nullptr, "couette_steps: scratch allocation failed");
        if (h->V_up.size() != (size_t)h->n_cells || std::memcmp(h->V_up.data(), V, h->n_cells * sizeof(double))) {
            MZ_HIP(hipMemcpy(v, V, h->n_cells * sizeof(double), hipMemcpyHostToDevice));
            h->V_up.assign(V, V + h->n_cells);
        }
        dV = v;
    }
    int r = mz_squash(h, 1);   // returns at once when the species is known to be contiguous
    if (r) return r;
    const double mass = masses[0];
    if (avg_inv > 0.0) {
        MZ_CHECK(mass > 0.0 && dt > 0.0, "couette_avg_steps: wall properties need mass > 0 and dt > 0");
        r = mz_surf_prealloc(h, 1);
        if (r) return r;
    }
    if (mode == 2 && equal_weight) {
        // slab-resident kernel; falls through to the graph when the state or
        // shape does not allow it or when the launch aborted (state restored)
        int ran = 0;
        r = mz_couette_slab(h, n_steps, step0, dt, dV, props_every, &ran, avg_inv, mass);
        if (r) return r;
        if (ran) {
            h->last_path = 2;
            return 0;
        }
        if (h->slab_info[3] == -1 || h->slab_info[3] == -2) {
            // the launch did not run or was undone: the replay reads the
            // device's contiguity flag again
            h->sp[0].contig_known = false;
            r = mz_squash(h, 1);
            if (r) return r;
        }
        mode = 1;
    }
    if (mode == 2) mode = 1;
    const bool graph_ok = mode == 1 && equal_weight && (props_every == 0 || props_every == 1);
    h->last_path = graph_ok ? 1 : 0;
    if (!graph_ok) {
        // eager launches (also the path for ntc! with weight splitting)
        MzTimed tm(h, "step", n_steps);
        for (int32_t t = 0; t < n_steps; ++t) {
            uint64_t step = step0 + (uint64_t)t;
            int do_props = props_every > 0 && ((t + 1) % props_every) == 0;
            if (equal_weight) {
                r = enqueue_step(h, dV, dt, do_props, step, nullptr, avg_inv, mass);
                if (r) return r;
            } else {
                r = mz_ntc(h, 1, 1, dt, V, 1e-16, 0, (step << 4) | 1u, 0, 1, h->n_cells);
                if (r) return r;
                r = mz_convect(h, 1, dt, (step << 4) | 2u, 0);
                if (r) return r;
                r = mz_sort(h, 1, 1);
                if (r) return r;
                if (do_props) {
                    r = mz_launch_props(h, 1, 0);
                    if (r) return r;
                }
            }
        }
        h->sp[0].contig_known = true;
        return 0;
    }
    // graph replay
    char keybuf[256];
    std::snprintf(keybuf, sizeof(keybuf), "couette:%a:%d:%p:%a:%a:%a:%a:%d", dt, props_every, (const void *)dV,
                  avg_inv, mass, h->surf_inv_area[0], h->surf_inv_area[1], h->sort_path);
    if (h->graph_key != keybuf) {
        mz_graph_release(h);
        MZ_HIP(hipMalloc(&h->d_step, sizeof(uint64_t)));
        // scratch is allocated before capture
        r = mz_sort_prealloc(h, 1);
        if (r) return r;
        MZ_HIP(hipStreamSynchronize(h->stream));
        MZ_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int rep = 0; rep < 2; ++rep) {
            r = enqueue_step(h, dV, dt, props_every == 1, 0, h->d_step, avg_inv, mass);
            if (r) {
                hipGraph_t g;
                hipStreamEndCapture(h->stream, &g);
                if (g) hipGraphDestroy(g);
                return r;
            }
            hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, h->stream, h->d_step);
        }
        MZ_HIP(hipStreamEndCapture(h->stream, &h->graph));
        MZ_HIP(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
        h->graph_key = keybuf;
    }
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, h->stream, h->d_step, step0);
    MzTimed tm(h, "step", n_steps);
    for (int32_t t = 0; t + 1 < n_steps; t += 2) MZ_HIP(hipGraphLaunch(h->graph_exec, h->stream));
    if (n_steps & 1) {
        uint64_t step = step0 + (uint64_t)(n_steps - 1);
        r = enqueue_step(h, dV, dt, props_every == 1, step, nullptr, avg_inv, mass);
        if (r) return r;
    }
    h->sp[0].contig_known = true;
    return 0;
}
