// C ABI of the projective smoothing path (declared in include/illcrawl_psg.h).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <memory>
#include <exception>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "gridbuild.hpp"
#include "kernels.hpp"

namespace ilc {
const char* last_error_cstr();
std::vector<float> cubic_projected_table(int n);
std::vector<float> cubic_projected_dtable(int n);
}  // namespace ilc

using namespace ilc;

struct ilc_psg {
  int device = 0;
  uint64_t target = 1000;
  int mode = ILC_MODE_REFERENCE;
  int kernel = ILC_KERNEL_QUARTIC_REF;
  tile_shape ts{32, 4};
  uint64_t blocksize = 0;

  grid_store grid;
  grid_scratch scratch;
  bbox_result* bbox_host = nullptr;
  uint32_t* flags_host = nullptr;

  // points
  const float4* pts = nullptr;
  uint64_t npts = 0;
  uint32_t lnx = 0, lny = 0;
  dbuf<double> acc;

  dbuf<float> cubic, dcubic;
  bool cubic_ready = false, dcubic_ready = false;
  dbuf<unsigned long long> cnt, cnt_scratch;
  dbuf<float> grad_sorted;  // adjoint: 4 x n gradients in grid order

  ~ilc_psg() {
    if (bbox_host) (void)hipHostFree(bbox_host);
    if (flags_host) (void)hipHostFree(flags_host);
  }
};

// dm::grid: the volumetric smoothing shares the single-view state
struct ilc_grid3d {
  ilc_psg p;
};

namespace {

template <class F>
ilc_status guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return ILC_OK;
  } catch (const ilc::error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return ILC_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ILC_EINVAL;
  }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

struct device_guard {
  int prev = -1;
  explicit device_guard(int dev) {
    ILC_HIP(hipGetDevice(&prev));
    if (prev != dev) ILC_HIP(hipSetDevice(dev));
  }
  ~device_guard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
  }
};

#define ILC_GUARD_NULL(h)                       \
  do {                                          \
    if (!(h)) {                                 \
      set_last_error("null handle");            \
      return ILC_EINVAL;                        \
    }                                           \
  } while (0)

void check_handle(const ilc_psg* h) { ILC_REQUIRE(h != nullptr, "null handle"); }

}  // namespace

extern "C" {

const char* ilc_last_error(void) { return ilc::last_error_cstr(); }

int ilc_abi_version(void) { return 1; }

ilc_status ilc_psg_create(const ilc_psg_config* cfg, ilc_psg** out) {
  return guarded([&] {
    ILC_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    ilc_psg_config c{};
    if (cfg) c = *cfg;
    ILC_REQUIRE(c.mode == ILC_MODE_REFERENCE || c.mode == ILC_MODE_EXACT, "bad mode %d", c.mode);
    ILC_REQUIRE(c.kernel >= ILC_KERNEL_QUARTIC_REF && c.kernel <= ILC_KERNEL_CUBIC,
                "bad kernel %d", c.kernel);
    ILC_REQUIRE(!(c.mode == ILC_MODE_REFERENCE && c.kernel == ILC_KERNEL_CUBIC),
                "reference mode uses a quartic kernel");
    int ndev = 0;
    ILC_HIP(hipGetDeviceCount(&ndev));
    ILC_REQUIRE(c.device >= 0 && c.device < ndev, "device %d not present (%d devices)", c.device,
                ndev);
    tile_shape ts{32, 4};
    if (c.tile_w || c.tile_h) {
      ILC_REQUIRE(c.tile_w > 0 && c.tile_h > 0 && (c.tile_h * c.tile_w) % kBlock == 0,
                  "tile %dx%d is not a multiple of %d points", c.tile_w, c.tile_h, kBlock);
      ts.tw = c.tile_w;
      ts.ppt = c.tile_w * c.tile_h / kBlock;
    }
    ilc_psg* h = new ilc_psg();
    h->device = c.device;
    h->target = c.target_particles_per_cell ? c.target_particles_per_cell : 1000;
    h->mode = c.mode;
    h->kernel = c.kernel;
    h->ts = ts;
    *out = h;
  });
}

ilc_status ilc_psg_destroy(ilc_psg* h) {
  return guarded([&] {
    if (!h) return;
    device_guard dg(h->device);
    delete h;
  });
}

const char* ilc_psg_backend_name(const ilc_psg* h) {
  (void)h;
  return "dm_smoothing/grid (hip gfx950)";
}

ilc_status ilc_psg_init_backend(ilc_psg* h, uint64_t blocksize) {
  return guarded([&] {
    check_handle(h);
    h->blocksize = blocksize;
  });
}

ilc_status ilc_psg_setup_projected_particles(ilc_psg* h, const float* particles4_dev, uint64_t n,
                                             const float* hsml_dev, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (h->blocksize == 0) fail(ILC_ENOTREADY, "init_backend was not called");
    ILC_REQUIRE(n > 0, "no particles");
    ILC_REQUIRE(n <= h->blocksize, "%llu particles exceed the blocksize %llu",
                (unsigned long long)n, (unsigned long long)h->blocksize);
    ILC_REQUIRE(n < (uint64_t(1) << 31), "block of %llu particles exceeds 2^31",
                (unsigned long long)n);
    ILC_REQUIRE(particles4_dev && hsml_dev, "null particle or smoothing-length pointer");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    const float4* p4 = reinterpret_cast<const float4*>(particles4_dev);
    if (!h->bbox_host) {
      ILC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->bbox_host), sizeof(bbox_result)));
      ILC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->flags_host), 2 * sizeof(uint32_t)));
    }
    grid_scratch& sc = h->scratch;
    // 1. bounding box (the host scan of particle_grid.cpp:43-67)
    const int nparts = static_cast<int>(std::min<uint64_t>(1024, (n + kBlock - 1) / kBlock));
    sc.bbox_partial.ensure(size_t(nparts) * 8);
    sc.bbox_dev.ensure(1);
    sc.flags.ensure(2);
    launch_bbox(p4, hsml_dev, n, sc.bbox_partial.p, nparts, sc.flags.p, sc.bbox_dev.p, s);
    ILC_HIP(hipMemcpyAsync(h->bbox_host, sc.bbox_dev.p, sizeof(bbox_result),
                           hipMemcpyDeviceToHost, s));
    ILC_HIP(hipStreamSynchronize(s));
    const bbox_result bb = *h->bbox_host;
    // 2. grid parameters (host, double, particle_grid.cpp:69-100)
    int32_t kz0 = 0;
    const grid_params g = plan_grid(bb, n, h->target, h->mode, &kz0);
    // 3. keys, stable sort, SoA, counts, max h, offsets
    launch_fill_u32(sc.flags.p, 0u, 1, s);
    enqueue_grid_build(h->grid, sc, p4, hsml_dev, n, g, kz0, bb.max_h, h->kernel, sc.flags.p, s);
    ILC_HIP(hipMemcpyAsync(h->flags_host, sc.flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ILC_HIP(hipStreamSynchronize(s));
    if (h->flags_host[0]) {
      h->grid.ready = false;
      fail(ILC_ERANGE, "%u particles fall outside the grid", h->flags_host[0]);
    }
  });
}

ilc_status ilc_psg_setup_evaluation_points(ilc_psg* h, const float* points4_dev, uint64_t n,
                                           uint32_t layout_nx, void* stream) {
  return guarded([&] {
    check_handle(h);
    ILC_REQUIRE(n > 0, "no evaluation points");
    ILC_REQUIRE(points4_dev, "null point pointer");
    ILC_REQUIRE(n < (uint64_t(1) << 32), "too many points");
    device_guard dg(h->device);
    h->pts = reinterpret_cast<const float4*>(points4_dev);
    h->npts = n;
    h->lnx = layout_nx ? layout_nx : uint32_t(h->ts.tw);
    h->lny = uint32_t((n + h->lnx - 1) / h->lnx);
    h->acc.ensure(n);
    launch_fill_f64(h->acc.p, 0.0, n, as_stream(stream));
  });
}

static void ensure_cubic(ilc_psg* h) {
  if (h->kernel != ILC_KERNEL_CUBIC || h->cubic_ready) return;
  std::vector<float> t = cubic_projected_table(4096);
  h->cubic.ensure(t.size());
  ILC_HIP(hipMemcpy(h->cubic.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  h->cubic_ready = true;
}

ilc_status ilc_psg_run(ilc_psg* h, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->grid.ready) fail(ILC_ENOTREADY, "run before setup_projected_particles");
    if (!h->pts) fail(ILC_ENOTREADY, "run before setup_evaluation_points");
    device_guard dg(h->device);
    ensure_cubic(h);
    point_view p{h->pts, static_cast<uint32_t>(h->npts), h->lnx, h->lny};
    launch_gather(h->grid.view(), p, h->mode == ILC_MODE_REFERENCE ? OP_REFERENCE : OP_EXACT,
                  h->kernel, h->ts, h->acc.p, nullptr, h->cubic.p, as_stream(stream));
  });
}

ilc_status ilc_psg_backward(ilc_psg* h, const float* dl_dvalue_dev, float scale, float* dl_dw_dev,
                            float* dl_dh_dev, float* dl_dx_dev, float* dl_dy_dev, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->grid.ready) fail(ILC_ENOTREADY, "backward before setup_projected_particles");
    if (!h->pts) fail(ILC_ENOTREADY, "backward before setup_evaluation_points");
    if (h->mode != ILC_MODE_EXACT)
      fail(ILC_EINVAL, "backward needs exact mode (the reference-mode search box is not smooth)");
    ILC_REQUIRE(dl_dvalue_dev && dl_dw_dev && dl_dh_dev && dl_dx_dev && dl_dy_dev,
                "null gradient pointer");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    const uint64_t n = h->grid.n;
    if (!n) return;
    h->grad_sorted.ensure(4 * n);
    float* gs = h->grad_sorted.p;
    launch_fill_u32(reinterpret_cast<uint32_t*>(gs), 0u, 4 * n, s);
    point_view p{h->pts, static_cast<uint32_t>(h->npts), h->lnx, h->lny};
    ensure_cubic(h);
    if (h->kernel == ILC_KERNEL_CUBIC && !h->dcubic_ready) {
      const std::vector<float> t = cubic_projected_dtable(4096);
      h->dcubic.ensure(t.size());
      ILC_HIP(hipMemcpy(h->dcubic.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
      h->dcubic_ready = true;
    }
    launch_adjoint(h->grid.view(), p, h->kernel, dl_dvalue_dev, scale, kernel_norm(h->kernel),
                   h->cubic.p, h->dcubic.p, gs, gs + n, gs + 2 * n, gs + 3 * n, s);
    launch_unpermute4(h->grid.perm.p, static_cast<uint32_t>(n), gs, gs + n, gs + 2 * n, gs + 3 * n,
                      dl_dw_dev, dl_dh_dev, dl_dx_dev, dl_dy_dev, s);
  });
}

ilc_status ilc_psg_retrieve_results(ilc_psg* h, float scale, float* out_dev, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->pts) fail(ILC_ENOTREADY, "no evaluation points");
    ILC_REQUIRE(out_dev, "null output");
    device_guard dg(h->device);
    launch_retrieve(h->acc.p, h->npts, scale, out_dev, as_stream(stream));
  });
}

ilc_status ilc_psg_retrieve_results_f64(ilc_psg* h, double* out_dev, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->pts) fail(ILC_ENOTREADY, "no evaluation points");
    ILC_REQUIRE(out_dev, "null output");
    device_guard dg(h->device);
    ILC_HIP(hipMemcpyAsync(out_dev, h->acc.p, sizeof(double) * h->npts, hipMemcpyDeviceToDevice,
                           as_stream(stream)));
  });
}

ilc_status ilc_psg_grid_info(const ilc_psg* h, ilc_grid_info* out) {
  return guarded([&] {
    check_handle(h);
    ILC_REQUIRE(out, "null output");
    if (!h->grid.ready) fail(ILC_ENOTREADY, "no grid");
    for (int a = 0; a < 3; ++a) {
      out->num_cells[a] = h->grid.gp.nc[a];
      out->min_corner[a] = h->grid.gp.minc[a];
      out->cell_size[a] = h->grid.gp.cs[a];
    }
    out->max_h = h->grid.max_h;
    out->num_particles = h->grid.n;
  });
}

ilc_status ilc_psg_export_grid(const ilc_psg* h, int32_t* cells2_dev, float* maxh_dev,
                               uint32_t* perm_dev, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->grid.ready) fail(ILC_ENOTREADY, "no grid");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    launch_export_cells(h->grid.start.p, h->grid.maxh_bits.p,
                        static_cast<uint32_t>(h->grid.gp.ncells()), cells2_dev, maxh_dev, s);
    if (perm_dev && h->grid.n) {
      // the gather keeps a cell's particles by descending h; the reference's
      // order within a cell is the input order
      const uint64_t n = h->grid.n;
      dbuf<uint32_t> by_index, temp;
      by_index.ensure(n);
      temp.ensure(radix_temp_words(n));
      launch_reference_perm(h->grid.key.p, h->grid.perm.p, static_cast<uint32_t>(n),
                            key_bits_for(h->grid.gp.ncells() - 1), perm_dev, by_index.p, temp.p, s);
      ILC_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
    }
  });
}

ilc_status ilc_psg_count_evals(ilc_psg* h, uint64_t* counts_dev, uint64_t* total, void* stream) {
  return guarded([&] {
    check_handle(h);
    if (!h->grid.ready) fail(ILC_ENOTREADY, "no grid");
    if (!h->pts) fail(ILC_ENOTREADY, "no evaluation points");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    h->cnt.ensure(h->npts);
    h->cnt_scratch.ensure(1);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->cnt.p), 0u, 2 * h->npts, s);
    point_view p{h->pts, static_cast<uint32_t>(h->npts), h->lnx, h->lny};
    launch_gather(h->grid.view(), p, OP_COUNT, h->kernel, h->ts, nullptr, h->cnt.p, nullptr, s);
    if (counts_dev)
      ILC_HIP(hipMemcpyAsync(counts_dev, h->cnt.p, sizeof(unsigned long long) * h->npts,
                             hipMemcpyDeviceToDevice, s));
    const unsigned long long tot = launch_sum_u64(h->cnt.p, h->npts, h->cnt_scratch.p, s);
    if (total) *total = tot;
  });
}

ilc_status ilc_camera_projection(const ilc_camera* cam, float minv[9], float center[3]) {
  return guarded([&] {
    ILC_REQUIRE(cam, "null camera");
    const camera_frame f = make_camera_frame(*cam);
    for (int i = 0; i < 9; ++i)
      if (minv) minv[i] = f.minv[i];
    for (int i = 0; i < 3; ++i)
      if (center) center[i] = f.center[i];
  });
}

ilc_status ilc_camera_basis(const ilc_camera* cam, double b0[3], double b1[3],
                            double min_position[3], double* pixel_size) {
  return guarded([&] {
    ILC_REQUIRE(cam, "null camera");
    const camera_frame f = make_camera_frame(*cam);
    for (int i = 0; i < 3; ++i) {
      if (b0) b0[i] = f.b0[i];
      if (b1) b1[i] = f.b1[i];
      if (min_position) min_position[i] = f.min_pos[i];
    }
    if (pixel_size) *pixel_size = f.pixel_size;
  });
}

ilc_status ilc_project_particles(const ilc_camera* cam, const float* in4_dev, uint64_t n,
                                 float* out4_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(cam && in4_dev && out4_dev, "null argument");
    const camera_frame f = make_camera_frame(*cam);
    launch_project_particles(f, static_cast<float>(cam->depth),
                             reinterpret_cast<const float4*>(in4_dev), n,
                             reinterpret_cast<float4*>(out4_dev), as_stream(stream));
  });
}

ilc_status ilc_camera_pixel_points(const ilc_camera* cam, float* out4_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(cam && out4_dev, "null argument");
    const camera_frame f = make_camera_frame(*cam);
    launch_pixel_points(f, cam->nx, cam->ny, reinterpret_cast<float4*>(out4_dev),
                        as_stream(stream));
  });
}

ilc_status ilc_camera_generate_pixel_coordinates(const ilc_camera* cam, float* out4_dev,
                                                 void* stream) {
  return guarded([&] {
    ILC_REQUIRE(cam && out4_dev, "null argument");
    const camera_frame f = make_camera_frame(*cam);
    launch_world_pixels(f, cam->nx, cam->ny, reinterpret_cast<float4*>(out4_dev),
                        as_stream(stream));
  });
}

ilc_status ilc_project_evaluation_points(const ilc_camera* cam, const float* in4_dev, uint64_t n,
                                         float* out4_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(cam && in4_dev && out4_dev, "null argument");
    const camera_frame f = make_camera_frame(*cam);
    launch_project_points(f, reinterpret_cast<const float4*>(in4_dev), n,
                          reinterpret_cast<float4*>(out4_dev), as_stream(stream));
  });
}

ilc_status ilc_vector_scale(float* x_dev, uint64_t n, float alpha, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(x_dev || n == 0, "null argument");
    launch_vector_scale(x_dev, alpha, n, as_stream(stream));
  });
}

// ------------------------------------------------ volumetric smoothing (dm::grid)

ilc_status ilc_grid3d_create(const ilc_psg_config* cfg, ilc_grid3d** out) {
  return guarded([&] {
    ILC_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    ilc_psg_config c{};
    if (cfg) c = *cfg;
    ILC_REQUIRE(c.mode == ILC_MODE_REFERENCE || c.mode == ILC_MODE_EXACT, "bad mode %d", c.mode);
    ILC_REQUIRE(c.kernel == ILC_KERNEL_QUARTIC_REF || c.kernel == ILC_KERNEL_QUARTIC,
                "the volumetric smoothing uses the quartic kernel (got %d)", c.kernel);
    int ndev = 0;
    ILC_HIP(hipGetDeviceCount(&ndev));
    ILC_REQUIRE(c.device >= 0 && c.device < ndev, "device %d not present (%d devices)", c.device,
                ndev);
    std::unique_ptr<ilc_grid3d> h(new ilc_grid3d());
    h->p.device = c.device;
    // dm::grid's default of 64 particles per cell (dm_reconstruction_backend_grid.hpp:39)
    h->p.target = c.target_particles_per_cell ? c.target_particles_per_cell : 64;
    h->p.mode = c.mode;
    h->p.kernel = c.kernel;
    h->p.ts = tile_shape{16, 1};  // the point layout's default width
    *out = h.release();
  });
}

ilc_status ilc_grid3d_destroy(ilc_grid3d* h) {
  return guarded([&] {
    if (!h) return;
    device_guard dg(h->p.device);
    delete h;
  });
}

ilc_status ilc_grid3d_init_backend(ilc_grid3d* h, uint64_t blocksize) {
  ILC_GUARD_NULL(h);
  return ilc_psg_init_backend(&h->p, blocksize);
}

ilc_status ilc_grid3d_setup_particles(ilc_grid3d* g3, const float* particles4_dev, uint64_t n,
                                      const float* hsml_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(g3 != nullptr, "null handle");
    ilc_psg* h = &g3->p;
    if (h->blocksize == 0) fail(ILC_ENOTREADY, "init_backend was not called");
    ILC_REQUIRE(n > 0, "no particles");
    ILC_REQUIRE(n <= h->blocksize, "%llu particles exceed the blocksize %llu",
                (unsigned long long)n, (unsigned long long)h->blocksize);
    ILC_REQUIRE(n < (uint64_t(1) << 31), "block of %llu particles exceeds 2^31",
                (unsigned long long)n);
    ILC_REQUIRE(particles4_dev && hsml_dev, "null particle or smoothing-length pointer");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    const float4* p4 = reinterpret_cast<const float4*>(particles4_dev);
    if (!h->bbox_host) {
      ILC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->bbox_host), sizeof(bbox_result)));
      ILC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->flags_host), 2 * sizeof(uint32_t)));
    }
    grid_scratch& sc = h->scratch;
    // bounding box (particle_grid.cpp:43-67), then the grid parameters
    const int nparts = static_cast<int>(std::min<uint64_t>(1024, (n + kBlock - 1) / kBlock));
    sc.bbox_partial.ensure(size_t(nparts) * 8);
    sc.bbox_dev.ensure(1);
    sc.flags.ensure(2);
    launch_bbox(p4, hsml_dev, n, sc.bbox_partial.p, nparts, sc.flags.p, sc.bbox_dev.p, s);
    ILC_HIP(hipMemcpyAsync(h->bbox_host, sc.bbox_dev.p, sizeof(bbox_result),
                           hipMemcpyDeviceToHost, s));
    ILC_HIP(hipStreamSynchronize(s));
    const bbox_result bb = *h->bbox_host;
    if (bb.bad)
      fail(ILC_EINVAL, "%u particles have non-finite coordinates/w or h outside [1.1e-19, 1.8e19]",
           bb.bad);
    const grid_params g = compute_grid_params(bb, n, h->target);
    // the masked list packs a cell's x and y into 16 bits each
    if (h->mode == ILC_MODE_REFERENCE && (g.nc[0] > 65536 || g.nc[1] > 65536))
      fail(ILC_ERANGE, "grid of %d x %d cells in x, y (at most 65536 each)", g.nc[0], g.nc[1]);
    launch_fill_u32(sc.flags.p, 0u, 1, s);
    enqueue_volumetric_grid(h->grid, sc, p4, hsml_dev, n, g, bb.max_h, h->kernel, sc.flags.p, s);
    ILC_HIP(hipMemcpyAsync(h->flags_host, sc.flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ILC_HIP(hipStreamSynchronize(s));
    if (h->flags_host[0]) {
      h->grid.ready = false;
      fail(ILC_ERANGE, "%u particles fall outside the grid", h->flags_host[0]);
    }
  });
}

ilc_status ilc_grid3d_setup_evaluation_points(ilc_grid3d* h, const float* points4_dev,
                                              uint64_t n, uint32_t layout_nx, void* stream) {
  ILC_GUARD_NULL(h);
  return ilc_psg_setup_evaluation_points(&h->p, points4_dev, n, layout_nx, stream);
}

ilc_status ilc_grid3d_run(ilc_grid3d* g3, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(g3 != nullptr, "null handle");
    ilc_psg* h = &g3->p;
    if (!h->grid.ready) fail(ILC_ENOTREADY, "run before setup_particles");
    if (!h->pts) fail(ILC_ENOTREADY, "run before setup_evaluation_points");
    device_guard dg(h->device);
    point_view p{h->pts, static_cast<uint32_t>(h->npts), h->lnx, h->lny};
    launch_smooth3d(h->grid.view(), h->grid.levels(), p,
                    h->mode == ILC_MODE_REFERENCE ? OP_REFERENCE : OP_EXACT,
                    h->acc.p, as_stream(stream));
  });
}

ilc_status ilc_grid3d_retrieve_results(ilc_grid3d* h, float scale, float* out_dev, void* stream) {
  ILC_GUARD_NULL(h);
  return ilc_psg_retrieve_results(&h->p, scale, out_dev, stream);
}

ilc_status ilc_grid3d_retrieve_results_f64(ilc_grid3d* h, double* out_dev, void* stream) {
  ILC_GUARD_NULL(h);
  return ilc_psg_retrieve_results_f64(&h->p, out_dev, stream);
}

ilc_status ilc_grid3d_grid_info(const ilc_grid3d* h, ilc_grid_info* out) {
  ILC_GUARD_NULL(h);
  return ilc_psg_grid_info(&h->p, out);
}

ilc_status ilc_grid3d_count_evals(ilc_grid3d* g3, uint64_t* counts_dev, uint64_t* total,
                                  void* stream) {
  return guarded([&] {
    ILC_REQUIRE(g3 != nullptr, "null handle");
    ilc_psg* h = &g3->p;
    if (!h->grid.ready) fail(ILC_ENOTREADY, "no grid");
    if (!h->pts) fail(ILC_ENOTREADY, "no evaluation points");
    device_guard dg(h->device);
    hipStream_t s = as_stream(stream);
    h->cnt.ensure(h->npts);
    h->cnt_scratch.ensure(1);
    launch_fill_u32(reinterpret_cast<uint32_t*>(h->cnt.p), 0u, 2 * h->npts, s);
    point_view p{h->pts, static_cast<uint32_t>(h->npts), h->lnx, h->lny};
    launch_smooth3d(h->grid.view(), h->grid.levels(), p, OP_COUNT, nullptr, s, h->cnt.p);
    if (counts_dev)
      ILC_HIP(hipMemcpyAsync(counts_dev, h->cnt.p, sizeof(unsigned long long) * h->npts,
                             hipMemcpyDeviceToDevice, s));
    const unsigned long long tot = launch_sum_u64(h->cnt.p, h->npts, h->cnt_scratch.p, s);
    if (total) *total = tot;
  });
}

ilc_status ilc_particles_set_w(float* particles4_dev, const float* w_dev, uint64_t n,
                               void* stream) {
  return guarded([&] {
    ILC_REQUIRE(n == 0 || (particles4_dev && w_dev), "null argument");
    launch_set_w(reinterpret_cast<float4*>(particles4_dev), w_dev, n, as_stream(stream));
  });
}

}  // extern "C"

namespace {
// Shared body of ilc_crawl_block and ilc_crawl_block_static: `get_scratch`
// returns the counts / offsets / scan scratch of the given number of words.
template <class Scratch>
void crawl_block_body(const ilc_volume* vol, const float* xyz_dev, uint64_t n,
                      const float* const* columns_dev, int32_t ncolumns, float* particles4_dev,
                      float* const* columns_out_dev, uint64_t* index_out_dev, uint64_t* n_out,
                      void* stream, Scratch&& get_scratch) {
  ILC_REQUIRE(vol && n_out, "null argument");
  ILC_REQUIRE(ncolumns >= 0 && ncolumns <= kMaxCrawlColumns, "ncolumns %d not in [0, %d]",
              ncolumns, kMaxCrawlColumns);
  ILC_REQUIRE(n == 0 || (xyz_dev && particles4_dev), "null argument");
  ILC_REQUIRE(ncolumns == 0 || (columns_dev && columns_out_dev), "null column list");
  *n_out = 0;
  if (!n) return;
  crawl_columns cols{};
  cols.n = ncolumns;
  for (int k = 0; k < ncolumns; ++k) {
    ILC_REQUIRE(columns_dev[k] && columns_out_dev[k], "null column %d", k);
    cols.in[k] = columns_dev[k];
    cols.out[k] = columns_out_dev[k];
  }
  double vmin[3], vmax[3], gmin[3], gmax[3];
  for (int j = 0; j < 3; ++j) {
    vmin[j] = vol->center[j] - 0.5 * vol->extent[j];
    vmax[j] = vol->center[j] + 0.5 * vol->extent[j];
    gmin[j] = vmin[j] - vol->border;
    gmax[j] = vmax[j] + vol->border;
  }
  hipStream_t s = as_stream(stream);
  const uint64_t nt = crawl_tiles(n);
  uint32_t* work = get_scratch(2 * (nt + 1) + scan_temp_words(nt + 1));
  uint32_t* counts = work;
  uint32_t* offsets = work + nt + 1;
  // every tile's count is written by the count pass; the scan's extra
  // element must read 0
  launch_fill_u32(counts + nt, 0u, 1, s);
  launch_crawl(vol->periodic_wraparound, gmin, gmax, vmin, vmax, xyz_dev, n, cols, counts,
               offsets, work + 2 * (nt + 1), reinterpret_cast<float4*>(particles4_dev),
               index_out_dev, s);
  uint32_t m = 0;
  ILC_HIP(hipMemcpyAsync(&m, offsets + nt, sizeof m, hipMemcpyDeviceToHost, s));
  ILC_HIP(hipStreamSynchronize(s));
  *n_out = m;
}
}  // namespace

// Scratch of ilc_crawl_block (tile counts, offsets, scan sums), leased from a
// per-process pool: a call takes a free buffer of its device (allocating, or
// growing one, only when none is large enough) and returns it when done, so
// repeated blocks neither allocate nor free (hipFree synchronises the whole
// device), and concurrent calls get different buffers.  The pool lives until
// the process ends.
struct crawl_pool {
  struct buf {
    int device;
    uint32_t* p;
    uint64_t words;
  };
  std::mutex m;
  std::vector<buf> free;
};

crawl_pool& the_crawl_pool() {
  static crawl_pool* pool = new crawl_pool;  // never destroyed
  return *pool;
}

struct crawl_lease {
  crawl_pool::buf b{-1, nullptr, 0};
  uint32_t* get(uint64_t words) {
    int dev = 0;
    ILC_HIP(hipGetDevice(&dev));
    crawl_pool& pool = the_crawl_pool();
    {
      std::lock_guard<std::mutex> lk(pool.m);
      for (size_t k = 0; k < pool.free.size(); ++k)
        if (pool.free[k].device == dev && pool.free[k].words >= words) {
          b = pool.free[k];
          pool.free.erase(pool.free.begin() + long(k));
          return b.p;
        }
      // the device's largest free buffer is too small: replace it
      for (size_t k = 0; k < pool.free.size(); ++k)
        if (pool.free[k].device == dev) {
          (void)hipFree(pool.free[k].p);
          pool.free.erase(pool.free.begin() + long(k));
          break;
        }
    }
    void* p = nullptr;
    ILC_HIP(hipMalloc(&p, sizeof(uint32_t) * words));
    b = crawl_pool::buf{dev, static_cast<uint32_t*>(p), words};
    return b.p;
  }
  ~crawl_lease() {
    if (!b.p) return;
    // on an error the block's kernels may still be running on the buffer
    if (std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
    crawl_pool& pool = the_crawl_pool();
    std::lock_guard<std::mutex> lk(pool.m);
    pool.free.push_back(b);
  }
};

extern "C" {

ilc_status ilc_crawl_block(const ilc_volume* vol, const float* xyz_dev, uint64_t n,
                           const float* const* columns_dev, int32_t ncolumns,
                           float* particles4_dev, float* const* columns_out_dev,
                           uint64_t* index_out_dev, uint64_t* n_out, void* stream) {
  return guarded([&] {
    crawl_lease lease;
    crawl_block_body(vol, xyz_dev, n, columns_dev, ncolumns, particles4_dev, columns_out_dev,
                     index_out_dev, n_out, stream,
                     [&](uint64_t words) { return lease.get(words); });
  });
}

ilc_status ilc_crawl_block_static(const ilc_volume* vol, const float* xyz_dev, uint64_t n,
                                  const float* const* columns_dev, int32_t ncolumns,
                                  float* particles4_dev, float* const* columns_out_dev,
                                  uint64_t* index_out_dev, uint64_t* n_out, void* stream) {
  return guarded([&] {
    if (n > kCrawlStaticMaxRows)
      fail(ILC_ERANGE, "block of %llu rows exceeds the allocation-free maximum of %llu rows",
           (unsigned long long)n, (unsigned long long)kCrawlStaticMaxRows);
    static_assert(ILC_CRAWL_STATIC_MAX_ROWS == kCrawlStaticMaxRows, "header limit");
    crawl_block_body(vol, xyz_dev, n, columns_dev, ncolumns, particles4_dev, columns_out_dev,
                     index_out_dev, n_out, stream,
                     [](uint64_t words) { return crawl_static_scratch(words); });
  });
}

namespace {
// the Gaunt table on the device for the duration of one call
struct device_table {
  float* p = nullptr;
  device_table(const float* host, int32_t nx, int32_t ny, hipStream_t s) {
    ILC_REQUIRE(host, "null Gaunt table");
    ILC_REQUIRE(nx >= 1 && ny >= 1 && int64_t(nx) * ny <= (int64_t(1) << 24),
                "Gaunt table %d x %d out of range", nx, ny);
    ILC_HIP(hipMalloc(&p, sizeof(float) * size_t(nx) * ny));
    ILC_HIP(hipMemcpyAsync(p, host, sizeof(float) * size_t(nx) * ny, hipMemcpyHostToDevice, s));
  }
  ~device_table() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

ilc_status ilc_quantity_xray_flux(const float* rho_dev, const float* u_dev, const float* xe_dev,
                                  uint64_t n, const float* gaunt_host, int32_t gnx, int32_t gny,
                                  float z, float lum_dist, float e_min, float e_max,
                                  int32_t num_samples, float* w_dev, int32_t* steps_dev,
                                  void* stream) {
  return guarded([&] {
    ILC_REQUIRE(num_samples >= 1, "num_samples must be >= 1");
    if (!n) return;
    ILC_REQUIRE(rho_dev && u_dev && xe_dev && w_dev, "null argument");
    hipStream_t s = as_stream(stream);
    device_table t(gaunt_host, gnx, gny, s);
    launch_xray_flux(rho_dev, u_dev, xe_dev, n, t.p, gnx, gny, z, lum_dist, e_min, e_max,
                     num_samples, w_dev, steps_dev, s);
    ILC_HIP(hipStreamSynchronize(s));
  });
}

ilc_status ilc_quantity_xray_spectral_flux(const float* rho_dev, const float* u_dev,
                                           const float* xe_dev, uint64_t n,
                                           const float* gaunt_host, int32_t gnx, int32_t gny,
                                           float z, float lum_dist, float e, float de,
                                           float* w_dev, void* stream) {
  return guarded([&] {
    if (!n) return;
    ILC_REQUIRE(rho_dev && u_dev && xe_dev && w_dev, "null argument");
    hipStream_t s = as_stream(stream);
    device_table t(gaunt_host, gnx, gny, s);
    launch_xray_spectral_flux(rho_dev, u_dev, xe_dev, n, t.p, gnx, gny, z, lum_dist, e, de,
                              w_dev, s);
    ILC_HIP(hipStreamSynchronize(s));
  });
}

ilc_status ilc_quantity_scaled(const float* in_dev, double scale, uint64_t n, float* w_dev,
                               void* stream) {
  return guarded([&] {
    ILC_REQUIRE(w_dev, "null output");
    launch_quantity_scaled(in_dev, scale, n, w_dev, as_stream(stream));
  });
}

ilc_status ilc_quantity_stellar_mass(const float* formation_time_dev, const float* mass_dev,
                                     uint64_t n, float* w_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(w_dev && (n == 0 || (formation_time_dev && mass_dev)), "null argument");
    launch_quantity_stellar_mass(formation_time_dev, mass_dev, n, w_dev, as_stream(stream));
  });
}

ilc_status ilc_quantity_xray_bolometric(const float* rho_dev, const float* temp_dev,
                                        const float* xe_dev, uint64_t n,
                                        const float* table_host, int32_t ntab, float x0,
                                        float dx, float* w_dev, void* stream) {
  return guarded([&] {
    ILC_REQUIRE(rho_dev && temp_dev && xe_dev && w_dev && table_host, "null argument");
    ILC_REQUIRE(dx > 0.0f, "table spacing must be positive");
    ILC_REQUIRE(ntab >= 1 && ntab <= 4096, "table size %d not in [1, 4096]", ntab);
    hipStream_t s = as_stream(stream);
    static thread_local dbuf<float> table;
    table.ensure(size_t(ntab));
    ILC_HIP(hipMemcpyAsync(table.p, table_host, sizeof(float) * ntab, hipMemcpyHostToDevice, s));
    launch_xray_bolometric(rho_dev, temp_dev, xe_dev, n, table.p, ntab, x0, dx, w_dev, s);
    ILC_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
