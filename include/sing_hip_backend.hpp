// include/sing_hip_backend.hpp — HipBackend : singularity::RenderBackend, the
// reference's backend interface (render/include/render_backend.hpp:107-132)
// implemented over the C ABI of include/sing_abi.h, with CudaBackend's
// semantics (render/cuda/cuda_backend.cu:105-166):
//   * initialize() allocates the frame for config.width x config.height on
//     device 0 and returns false when that fails (cuda_backend.cu:111-118);
//   * render_frame() packs the Scene and camera into the 160-byte uniform block
//     (pack_uniforms_cuda, cuda_backend.cu:45-101), renders the reference's FP32
//     RGBA8 frame synchronously and advances the 1/30 s virtual clock
//     (cuda_backend.cu:149-151); without a frame it does nothing
//     (cuda_backend.cu:141-143); errors are printed to stderr and ignored, as
//     the reference does (cuda_backend.cu:38-43);
//   * capture_frame() copies the last frame out by value;
//   * name() is "HIP-gfx950".
// Header-only: a maintainer adds it next to render/include/render_backend.hpp and
// links libsing_hip.so.  Precision and extra outputs stay at the reference's
// defaults here; use the C ABI for the mixed-precision and HDR paths.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>

#include "render_backend.hpp"
#include "sing_abi.h"

namespace singularity {

class HipBackend final : public RenderBackend {
public:
    HipBackend() = default;
    ~HipBackend() override { shutdown(); }

    bool initialize(WindowHandle /*window: offline, ignored*/, RenderConfig config) override {
        shutdown();
        ctx_ = sing_create(config.width, config.height, /*device=*/0);
        if (!ctx_) {
            std::fprintf(stderr, "[hip] initialize failed: %s\n", sing_last_error(nullptr));
            return false;
        }
        width_ = config.width;
        height_ = config.height;
        frame_time_ = 0.0f;
        return true;
    }

    void shutdown() override {
        if (ctx_) sing_destroy(ctx_);
        ctx_ = nullptr;
    }

    void resize(std::uint32_t width, std::uint32_t height) override {
        if (!ctx_) return;
        if (sing_resize(ctx_, width, height) != SING_OK) {
            std::fprintf(stderr, "[hip] resize failed: %s\n", sing_last_error(ctx_));
            return;
        }
        width_ = width;
        height_ = height;
    }

    void render_frame(const Scene& scene, const CameraState& camera) override {
        if (!ctx_) return;
        SingScene s;
        sing_scene_defaults(&s);
        s.metric = scene.metric == Scene::MetricType::Kerr ? SING_METRIC_KERR : SING_METRIC_SCHWARZSCHILD;
        s.mass = scene.mass_solar;
        s.spin_a_over_M = scene.spin_a_over_M;
        s.disc_inner_M = scene.disc_inner_M;
        s.disc_outer_M = scene.disc_outer_M;
        s.doppler_on = scene.disc_doppler_on ? 1u : 0u;
        s.redshift_on = scene.disc_redshift_on ? 1u : 0u;
        s.disc_on = scene.disc_texture_on ? 1u : 0u;
        s.supersample = scene.render_supersample;
        s.exposure = scene.exposure;
        s.bloom_threshold = scene.bloom_threshold;
        s.bloom_strength = scene.bloom_strength;
        s.disc_turbulence = scene.disc_turbulence;
        s.disc_peak_T_K = scene.disc_peak_T_K;
        s.max_steps = scene.max_steps;
        s.h_step = scene.h_step;
        SingCamera cam;
        for (int i = 0; i < 3; ++i) cam.position[i] = camera.position[i];
        for (int i = 0; i < 9; ++i) cam.basis[i] = camera.basis[i];
        cam.fov_y_radians = camera.fov_y_radians;
        SingUniforms u;
        sing_pack_uniforms(&s, &cam, width_, height_, frame_time_, &u);
        const SingRenderOpts opts{};  // FP32, RGBA8: the reference kernel's frame
        if (sing_render(ctx_, &u, &opts, nullptr) != SING_OK)
            std::fprintf(stderr, "[hip] render failed: %s\n", sing_last_error(ctx_));
        frame_time_ += 1.0f / 30.0f;
    }

    ImageData capture_frame() override {
        ImageData img;
        if (!ctx_) return img;
        img.width = width_;
        img.height = height_;
        img.pixels_rgba.resize(std::size_t(width_) * height_ * 4);
        if (sing_capture_rgba8(ctx_, img.pixels_rgba.data(), img.pixels_rgba.size()) != SING_OK) {
            std::fprintf(stderr, "[hip] capture failed: %s\n", sing_last_error(ctx_));
            img.pixels_rgba.clear();
        }
        return img;
    }

    const char* name() const override { return sing_backend_name(); }

    // The virtual clock (seconds) the next render_frame packs.
    float frame_time() const { return frame_time_; }

private:
    SingContext* ctx_ = nullptr;
    std::uint32_t width_ = 0, height_ = 0;
    float frame_time_ = 0.0f;
};

inline std::unique_ptr<RenderBackend> create_hip_backend() { return std::make_unique<HipBackend>(); }

}  // namespace singularity
