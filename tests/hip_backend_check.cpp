// tests/hip_backend_check.cpp — drives include/sing_hip_backend.hpp through the
// reference's RenderBackend interface (render/include/render_backend.hpp),
// compiled against that header where the reference tree is present.
//
// Usage: hip_backend_check OUT.rgba W H SPP
// Renders two frames of the reference's Scene{} defaults with the cuda_cli
// orbital camera (30 M, 0.15 rad, azimuth 0, FOV 60), writes the second frame's
// RGBA8 to OUT.rgba and prints "name=<backend> t=<clock>".  Exit 3 when
// initialize() fails (no device), 2 on usage errors.
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "sing_hip_backend.hpp"

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s OUT.rgba W H SPP\n", argv[0]);
        return 2;
    }
    std::unique_ptr<singularity::RenderBackend> be = singularity::create_hip_backend();
    singularity::RenderConfig cfg;
    cfg.width = std::uint32_t(std::atoi(argv[2]));
    cfg.height = std::uint32_t(std::atoi(argv[3]));
    std::printf("name=%s\n", be->name());
    if (!be->initialize(singularity::WindowHandle{}, cfg)) return 3;
    singularity::Scene scene;
    scene.render_supersample = std::uint32_t(std::atoi(argv[4]));
    SingCamera c;
    sing_orbital_camera(30.0f, 0.15f, 0.0f, 60.0f, &c);
    singularity::CameraState cam;
    for (int i = 0; i < 3; ++i) cam.position[i] = c.position[i];
    for (int i = 0; i < 9; ++i) cam.basis[i] = c.basis[i];
    cam.fov_y_radians = c.fov_y_radians;
    be->render_frame(scene, cam);
    be->render_frame(scene, cam);
    const singularity::ImageData img = be->capture_frame();
    std::printf("t=%.9g size=%ux%u\n", double(static_cast<singularity::HipBackend*>(be.get())->frame_time()),
                img.width, img.height);
    FILE* f = std::fopen(argv[1], "wb");
    if (!f || std::fwrite(img.pixels_rgba.data(), 1, img.pixels_rgba.size(), f) != img.pixels_rgba.size()) return 1;
    std::fclose(f);
    be->shutdown();
    return img.pixels_rgba.empty() ? 1 : 0;
}
