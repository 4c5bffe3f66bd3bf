// wavefront.hip — film accumulation, camera set-up and the wavefront integrator behind pe_render.
#include "photon_hip.h"
#include "pe_internal.h"
#include "pe_math.h"

#include <cmath>
#include <string>

namespace pe {

__global__ void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// Clears a film with a fill kernel (no memset node, so the call is safe inside a captured graph).
int film_clear(pe_film* f) {
    const size_t px = static_cast<size_t>(f->width) * f->height;
    const size_t words[4] = {3 * px, px, px, 2};
    void* bufs[4] = {f->sum.get(), f->count.get(), f->lum2.get(), f->rejected.get()};
    for (int k = 0; k < 4; ++k) {
        if (!bufs[k] || words[k] == 0) continue;
        const int grid = static_cast<int>(std::min<size_t>((words[k] + 255) / 256, 4096));
        hipLaunchKernelGGL(k_zero_u32, dim3(grid), dim3(256), 0, f->dev->stream, static_cast<uint32_t*>(bufs[k]),
                           words[k]);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("film clear: ") + hipGetErrorString(e));
        return PE_ERR_HIP;
    }
    return PE_OK;
}

}  // namespace pe

using namespace pe;

extern "C" {

// PerspectiveCamera constructor (camera/perspective_camera.cpp:11-34) and describeGpu (:42-54).
int pe_camera_perspective(const float position[3], const float target[3], const float up[3], float vfov_degrees,
                          float aspect, pe_camera* out) {
#pragma clang fp contract(off)
    if (!position || !target || !up || !out) {
        set_error("pe_camera_perspective: NULL argument");
        return PE_ERR_INVALID;
    }
    const V3 pos = mk(position[0], position[1], position[2]);
    const V3 tgt = mk(target[0], target[1], target[2]);
    const V3 upv = mk(up[0], up[1], up[2]);
    const float theta = vfov_degrees * (kPi / 180.0f);  // deg2rad (utils.h:44-46)
    const float h = std::tan(theta * 0.5f);
    const float vh = 2.0f * h;
    const float vw = aspect * vh;
    const V3 w = normalized(pos - tgt);
    const V3 u = normalized(cross(upv, w));
    const V3 v = cross(w, u);
    const V3 hor = vw * u;
    const V3 ver = vh * v;
    const V3 ll = pos - hor * 0.5f - ver * 0.5f - w;
    *out = pe_camera{};
    const V3 vs[7] = {pos, ll, hor, ver, u, v, -w};
    float* dst[7] = {out->position, out->lower_left, out->horizontal, out->vertical, out->u, out->v, out->forward};
    for (int k = 0; k < 7; ++k) {
        dst[k][0] = vs[k].x;
        dst[k][1] = vs[k].y;
        dst[k][2] = vs[k].z;
    }
    return PE_OK;
}

// ThinLensCamera constructor (camera/thin_lens_camera.cpp:14-33), setApertureShape (thin_lens_camera.h:31-34)
// and describeGpu (:63-77): the pinhole's image plane moved out to the focus distance.
int pe_camera_thin_lens(const float position[3], const float target[3], const float up[3], float vfov_degrees,
                        float aspect, float aperture_radius, float focus_distance, int blades, float rotation_degrees,
                        pe_camera* out) {
#pragma clang fp contract(off)
    if (!position || !target || !up || !out) {
        set_error("pe_camera_thin_lens: NULL argument");
        return PE_ERR_INVALID;
    }
    const V3 pos = mk(position[0], position[1], position[2]);
    const V3 tgt = mk(target[0], target[1], target[2]);
    const V3 upv = mk(up[0], up[1], up[2]);
    const float theta = vfov_degrees * (kPi / 180.0f);
    const float h = std::tan(theta * 0.5f);
    const float vh = 2.0f * h * focus_distance;
    const float vw = aspect * vh;
    const V3 w = normalized(pos - tgt);
    const V3 u = normalized(cross(upv, w));
    const V3 v = cross(w, u);
    const V3 hor = vw * u;
    const V3 ver = vh * v;
    const V3 ll = pos - hor * 0.5f - ver * 0.5f - focus_distance * w;
    *out = pe_camera{};
    const V3 vs[7] = {pos, ll, hor, ver, u, v, -w};
    float* dst[7] = {out->position, out->lower_left, out->horizontal, out->vertical, out->u, out->v, out->forward};
    for (int k = 0; k < 7; ++k) {
        dst[k][0] = vs[k].x;
        dst[k][1] = vs[k].y;
        dst[k][2] = vs[k].z;
    }
    out->lens_radius = aperture_radius;
    out->blades = blades >= 3 ? blades : 0;
    out->blade_rotation = rotation_degrees * 0.017453292519943295f;
    return PE_OK;
}

// OrthographicCamera constructor and describeGpu (camera/orthographic_camera.cpp:11-40): parallel rays along
// the view direction from a verticalExtent-high image rectangle through the camera position.
int pe_camera_orthographic(const float position[3], const float target[3], const float up[3], float vertical_extent,
                           float aspect, pe_camera* out) {
#pragma clang fp contract(off)
    if (!position || !target || !up || !out) {
        set_error("pe_camera_orthographic: NULL argument");
        return PE_ERR_INVALID;
    }
    const V3 pos = mk(position[0], position[1], position[2]);
    const V3 tgt = mk(target[0], target[1], target[2]);
    const V3 upv = mk(up[0], up[1], up[2]);
    const V3 w = normalized(pos - tgt);
    const V3 u = normalized(cross(upv, w));
    const V3 v = cross(w, u);
    const float height = vertical_extent > 1e-4f ? vertical_extent : 1.0f;
    const float width = aspect * height;
    const V3 hor = width * u;
    const V3 ver = height * v;
    const V3 ll = pos - hor * 0.5f - ver * 0.5f;
    const V3 centre = ll + hor * 0.5f + ver * 0.5f;
    const V3 du = dot(hor, hor) > 0.0f ? normalized(hor) : mk(1.0f, 0.0f, 0.0f);
    const V3 dv = dot(ver, ver) > 0.0f ? normalized(ver) : mk(0.0f, 1.0f, 0.0f);
    *out = pe_camera{};
    const V3 vs[7] = {centre, ll, hor, ver, du, dv, -w};
    float* dst[7] = {out->position, out->lower_left, out->horizontal, out->vertical, out->u, out->v, out->forward};
    for (int k = 0; k < 7; ++k) {
        dst[k][0] = vs[k].x;
        dst[k][1] = vs[k].y;
        dst[k][2] = vs[k].z;
    }
    out->orthographic = 1;
    return PE_OK;
}

// setShift (camera/perspective_camera.h:25-27, thin_lens_camera.h:25-27): lowerLeft += u (shiftX |vertical|) +
// vertical shiftY.
int pe_camera_set_shift(pe_camera* camera, float shift_x, float shift_y) {
#pragma clang fp contract(off)
    if (!camera) {
        set_error("pe_camera_set_shift: NULL argument");
        return PE_ERR_INVALID;
    }
    if (camera->orthographic) {
        set_error("pe_camera_set_shift: an orthographic camera has no lens shift");
        return PE_ERR_INVALID;
    }
    const V3 ver = mk(camera->vertical[0], camera->vertical[1], camera->vertical[2]);
    const V3 u = mk(camera->u[0], camera->u[1], camera->u[2]);
    const V3 ll = mk(camera->lower_left[0], camera->lower_left[1], camera->lower_left[2]) +
                  u * (shift_x * length(ver)) + ver * shift_y;
    camera->lower_left[0] = ll.x;
    camera->lower_left[1] = ll.y;
    camera->lower_left[2] = ll.z;
    return PE_OK;
}

// Camera::setShutterEnd(const Camera&) (camera/camera.h:49-54): the end camera's description becomes the pose
// at shutter close.
int pe_camera_set_shutter_end(pe_camera* camera, const pe_camera* end) {
    if (!camera || !end) {
        set_error("pe_camera_set_shutter_end: NULL argument");
        return PE_ERR_INVALID;
    }
    const float* src[7] = {end->position, end->lower_left, end->horizontal, end->vertical, end->u, end->v, end->forward};
    float* dst[7] = {camera->end_position, camera->end_lower_left, camera->end_horizontal, camera->end_vertical,
                     camera->end_u, camera->end_v, camera->end_forward};
    for (int k = 0; k < 7; ++k)
        for (int c = 0; c < 3; ++c) dst[k][c] = src[k][c];
    camera->end_lens_radius = end->lens_radius;
    camera->has_shutter_end = 1;
    return PE_OK;
}

int pe_film_create(pe_device* dev, int width, int height, pe_film** out) {
    if (!dev || !out || width <= 0 || height <= 0) {
        set_error("pe_film_create: bad argument");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    auto f = std::make_unique<pe_film>();
    f->dev = dev;
    f->width = width;
    f->height = height;
    const size_t px = static_cast<size_t>(width) * height;
    if (!f->sum.alloc(px * 12) || !f->count.alloc(px * 4) || !f->lum2.alloc(px * 4) || !f->rejected.alloc(8))
        return PE_ERR_OOM;
    int rc = film_clear(f.get());
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    *out = f.release();
    return PE_OK;
}

int pe_film_destroy(pe_film* film) {
    if (!film) return PE_OK;
    (void)hipSetDevice(film->dev->ordinal);
    (void)hipStreamSynchronize(film->dev->stream);
    delete film;
    return PE_OK;
}

int pe_film_clear(pe_film* film) {
    if (!film) {
        set_error("pe_film_clear: NULL film");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(film->dev->ordinal));
    int rc = film_clear(film);
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(film->dev->stream));
    return PE_OK;
}

int pe_film_download(pe_film* film, float* sum_rgb, uint32_t* count, float* sum_lum2, uint64_t* rejected) {
    if (!film) {
        set_error("pe_film_download: NULL film");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(film->dev->ordinal));
    hipStream_t s = film->dev->stream;
    const size_t px = static_cast<size_t>(film->width) * film->height;
    if (sum_rgb) PE_HIP_CHECK(hipMemcpyAsync(sum_rgb, film->sum.get(), px * 12, hipMemcpyDeviceToHost, s));
    if (count) PE_HIP_CHECK(hipMemcpyAsync(count, film->count.get(), px * 4, hipMemcpyDeviceToHost, s));
    if (sum_lum2) PE_HIP_CHECK(hipMemcpyAsync(sum_lum2, film->lum2.get(), px * 4, hipMemcpyDeviceToHost, s));
    if (rejected) PE_HIP_CHECK(hipMemcpyAsync(rejected, film->rejected.get(), 8, hipMemcpyDeviceToHost, s));
    PE_HIP_CHECK(hipStreamSynchronize(s));
    return PE_OK;
}

namespace {
int crypto_clear(pe_crypto_film* f) {
    const size_t px = static_cast<size_t>(f->width) * f->height;
    const size_t words[4] = {PE_CRYPTO_SLOTS * px, PE_CRYPTO_SLOTS * px, px, 2};
    void* bufs[4] = {f->ids.get(), f->counts.get(), f->totals.get(), f->dropped.get()};
    for (int k = 0; k < 4; ++k) {
        const int grid = static_cast<int>(std::min<size_t>((words[k] + 255) / 256, 4096));
        hipLaunchKernelGGL(k_zero_u32, dim3(grid), dim3(256), 0, f->dev->stream, static_cast<uint32_t*>(bufs[k]),
                           words[k]);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("crypto film clear: ") + hipGetErrorString(e));
        return PE_ERR_HIP;
    }
    return PE_OK;
}
}  // namespace

int pe_crypto_film_create(pe_device* dev, int width, int height, pe_crypto_film** out) {
    if (!dev || !out || width <= 0 || height <= 0) {
        set_error("pe_crypto_film_create: bad argument");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    auto f = std::make_unique<pe_crypto_film>();
    f->dev = dev;
    f->width = width;
    f->height = height;
    const size_t px = static_cast<size_t>(width) * height;
    if (!f->ids.alloc(px * PE_CRYPTO_SLOTS * 4) || !f->counts.alloc(px * PE_CRYPTO_SLOTS * 4) ||
        !f->totals.alloc(px * 4) || !f->dropped.alloc(8))
        return PE_ERR_OOM;
    int rc = crypto_clear(f.get());
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    *out = f.release();
    return PE_OK;
}

int pe_crypto_film_destroy(pe_crypto_film* film) {
    if (!film) return PE_OK;
    (void)hipSetDevice(film->dev->ordinal);
    (void)hipStreamSynchronize(film->dev->stream);
    delete film;
    return PE_OK;
}

int pe_crypto_film_clear(pe_crypto_film* film) {
    if (!film) {
        set_error("pe_crypto_film_clear: NULL film");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(film->dev->ordinal));
    int rc = crypto_clear(film);
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(film->dev->stream));
    return PE_OK;
}

int pe_crypto_film_download(pe_crypto_film* film, float* ids, uint32_t* counts, uint32_t* totals, uint64_t* dropped) {
    if (!film) {
        set_error("pe_crypto_film_download: NULL film");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(film->dev->ordinal));
    hipStream_t s = film->dev->stream;
    const size_t px = static_cast<size_t>(film->width) * film->height;
    if (ids) PE_HIP_CHECK(hipMemcpyAsync(ids, film->ids.get(), px * PE_CRYPTO_SLOTS * 4, hipMemcpyDeviceToHost, s));
    if (counts) PE_HIP_CHECK(hipMemcpyAsync(counts, film->counts.get(), px * PE_CRYPTO_SLOTS * 4, hipMemcpyDeviceToHost, s));
    if (totals) PE_HIP_CHECK(hipMemcpyAsync(totals, film->totals.get(), px * 4, hipMemcpyDeviceToHost, s));
    if (dropped) PE_HIP_CHECK(hipMemcpyAsync(dropped, film->dropped.get(), 8, hipMemcpyDeviceToHost, s));
    PE_HIP_CHECK(hipStreamSynchronize(s));
    return PE_OK;
}

}  // extern "C"
