// singularity_amd/csrc/hip_cli.cpp — offline renderer CLI over the C ABI.
//
// Flag-compatible with the reference's singularity_cuda_cli
// (cuda_cli/main.cpp:29-64,113-341): same flags, defaults, validation
// messages, exit codes (2 = usage/scene error, 1 = runtime failure), "wrote ..."
// lines, 2*pi multi-frame orbit and the 1/30 s virtual clock
// (render/cuda/cuda_backend.cu:149-151).  Additional flags select precision,
// extra outputs (HDR PFM, class-map PGM), the bloom post pass and the
// integrator settings the reference hard-codes.  PNGs are written by the
// encoder below (stored deflate blocks, CRC-32, Adler-32).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/sing_abi.h"

namespace {

// ---- PNG (RGBA8, filter 0, zlib stream of stored blocks) -----------------------
std::uint32_t crc32_table[256];
void crc_init() {
    for (std::uint32_t n = 0; n < 256; ++n) {
        std::uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : (c >> 1);
        crc32_table[n] = c;
    }
}
std::uint32_t crc32(const std::uint8_t* p, std::size_t n, std::uint32_t c = 0xFFFFFFFFu) {
    for (std::size_t i = 0; i < n; ++i) c = crc32_table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}
void put_be32(std::vector<std::uint8_t>& v, std::uint32_t x) {
    v.push_back(std::uint8_t(x >> 24));
    v.push_back(std::uint8_t(x >> 16));
    v.push_back(std::uint8_t(x >> 8));
    v.push_back(std::uint8_t(x));
}
void chunk(std::vector<std::uint8_t>& png, const char* type, const std::vector<std::uint8_t>& data) {
    put_be32(png, std::uint32_t(data.size()));
    const std::size_t at = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    put_be32(png, crc32(png.data() + at, data.size() + 4) ^ 0xFFFFFFFFu);
}
bool write_png(const char* path, const std::uint8_t* rgba, std::uint32_t w, std::uint32_t h) {
    crc_init();
    std::vector<std::uint8_t> raw;
    raw.reserve((std::size_t(w) * 4 + 1) * h);
    for (std::uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + std::size_t(y) * w * 4, rgba + std::size_t(y + 1) * w * 4);
    }
    std::vector<std::uint8_t> z = {0x78, 0x01};
    std::size_t pos = 0;
    do {
        const std::size_t n = std::min<std::size_t>(65535, raw.size() - pos);
        const bool final_block = pos + n == raw.size();
        z.push_back(final_block ? 1 : 0);
        z.push_back(std::uint8_t(n));
        z.push_back(std::uint8_t(n >> 8));
        z.push_back(std::uint8_t(~n));
        z.push_back(std::uint8_t(~n >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    std::uint32_t a = 1, b = 0;
    for (std::uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<std::uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<std::uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return std::fclose(f) == 0 && ok;
}
// Portable float map, RGB, little endian, rows bottom to top.
bool write_pfm(const char* path, const float* rgb, std::uint32_t w, std::uint32_t h) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    std::fprintf(f, "PF\n%u %u\n-1.0\n", w, h);
    bool ok = true;
    for (std::uint32_t y = h; y-- > 0;) ok &= std::fwrite(rgb + std::size_t(y) * w * 3, 4, std::size_t(w) * 3, f) == std::size_t(w) * 3;
    return std::fclose(f) == 0 && ok;
}
bool write_pgm(const char* path, const std::uint8_t* v, std::uint32_t w, std::uint32_t h) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    std::fprintf(f, "P5\n%u %u\n255\n", w, h);
    const bool ok = std::fwrite(v, 1, std::size_t(w) * h, f) == std::size_t(w) * h;
    return std::fclose(f) == 0 && ok;
}

void print_help() {
    std::puts("hip_cli — offline HIP renderer (MI355X / gfx950)\n"
              "\n"
              "Usage:\n"
              "  hip_cli [--scene PATH] [--output PATH] [--res WxH]\n"
              "          [--samples-per-pixel N]\n"
              "          [--camera-distance M] [--camera-elevation R]\n"
              "          [--camera-azimuth R] [--fov-y-deg D]\n"
              "          [--frames N --output-pattern PATTERN]\n"
              "\n"
              "Options (as singularity_cuda_cli):\n"
              "  --scene PATH         Scene config (key = value file).\n"
              "  --output PATH        PNG output path (default: cuda_render.png).\n"
              "  --res WxH            Output resolution (default: 1280x720)\n"
              "  --samples-per-pixel N  Halton(2,3) subpixel samples (1..1024, default 1).\n"
              "  --camera-distance M  Orbital radius in M (default 30)\n"
              "  --camera-elevation R Elevation from equator in rad (default 0.15)\n"
              "  --camera-azimuth R   Starting azimuth in rad (default 0.0)\n"
              "  --fov-y-deg D        Vertical FOV in degrees (default 60)\n"
              "  --frames N           Render N frames orbiting 2pi in azimuth (default 1).\n"
              "  --output-pattern P   printf-style path with one %d placeholder (required when --frames > 1).\n"
              "  --encode-mp4 PATH    After the last frame, run ffmpeg to encode the sequence.\n"
              "  --mp4-fps N          Encoder framerate (default 30).\n"
              "\n"
              "Additional options:\n"
              "  --precision P        fp32 | fp64 | mixed (default fp32)\n"
              "  --hdr-output PATH    Also write the HDR radiance (float RGB PFM).\n"
              "  --class-output PATH  Also write the termination-class map (PGM: 0 horizon, 1 escape,\n"
              "                       2 disc, 3 budget, 4 init failure).\n"
              "  --bloom              Bloom + ACES composite post pass for the PNG.\n"
              "  --h-step H           RK4 step in M (default 0.4)\n"
              "  --max-steps N        Step budget per ray (default 2500)\n"
              "  --escape-radius M    Escape radius in M (default 200)\n"
              "  --peak-temp-scale S  Multiplies the disc peak temperature (default 1)\n"
              "  --turb-seed N        Seeded turbulence band phases (0 = reference bands)\n"
              "  --adapt-tol T        Error-controlled RK4 with per-step tolerance T (Kerr scenes)\n"
              "  --sampler S          halton (default: Halton offsets, mean of tone-mapped samples) |\n"
              "                       grid (NxN grid, N = min(spp, 4), tone map of the mean radiance)\n"
              "  --refine LIST        Fixed-step Kerr refinements, comma-separated: cross64 (disc crossing\n"
              "                       at the FP64 dense-output root), reproject (null-constraint\n"
              "                       re-projection after every step)\n"
              "  --stats              Print per-frame timing and step counts.\n");
}

}  // namespace

int main(int argc, char** argv) {
    std::string scene_path, output_path = "cuda_render.png", output_pattern, encode_mp4_path, hdr_path, cls_path;
    std::uint32_t width = 1280, height = 720, spp = 1, frames = 1, mp4_fps = 30;
    float camera_distance = 30.0f, camera_elevation = 0.15f, camera_azimuth = 0.0f, fov_y_deg = 60.0f;
    float h_step = 0.0f, escape_r = 0.0f, temp_scale = 0.0f;
    std::uint32_t max_steps = 0, turb_seed = 0;
    SingRenderOpts opts{};
    bool bloom = false, print_stats = false;

    for (int i = 1; i < argc; ++i) {
        const std::string_view a = argv[i];
        const bool has = i + 1 < argc;
        if (a == "--help" || a == "-h") {
            print_help();
            return 0;
        }
        if (a == "--scene" && has) scene_path = argv[++i];
        else if (a == "--output" && has) output_path = argv[++i];
        else if (a == "--res" && has) {
            const char* v = argv[++i];
            const char* x = std::strchr(v, 'x');
            if (!x) {
                std::fprintf(stderr, "hip_cli: --res expects WxH, got %s\n", v);
                return 2;
            }
            width = std::uint32_t(std::strtoul(v, nullptr, 10));
            height = std::uint32_t(std::strtoul(x + 1, nullptr, 10));
            if (width < 4 || height < 4) {
                std::fprintf(stderr, "hip_cli: --res too small\n");
                return 2;
            }
        } else if (a == "--camera-distance" && has) camera_distance = std::strtof(argv[++i], nullptr);
        else if (a == "--camera-elevation" && has) camera_elevation = std::strtof(argv[++i], nullptr);
        else if (a == "--camera-azimuth" && has) camera_azimuth = std::strtof(argv[++i], nullptr);
        else if (a == "--fov-y-deg" && has) fov_y_deg = std::strtof(argv[++i], nullptr);
        else if (a == "--samples-per-pixel" && has) {
            const unsigned long v = std::strtoul(argv[++i], nullptr, 10);
            if (v < 1ul || v > 1024ul) {
                std::fprintf(stderr, "hip_cli: --samples-per-pixel must be in [1, 1024]\n");
                return 2;
            }
            spp = std::uint32_t(v);
        } else if (a == "--frames" && has) {
            const unsigned long v = std::strtoul(argv[++i], nullptr, 10);
            if (v < 1ul || v > 100000ul) {
                std::fprintf(stderr, "hip_cli: --frames must be in [1, 100000]\n");
                return 2;
            }
            frames = std::uint32_t(v);
        } else if (a == "--output-pattern" && has) output_pattern = argv[++i];
        else if (a == "--encode-mp4" && has) encode_mp4_path = argv[++i];
        else if (a == "--mp4-fps" && has) {
            const unsigned long v = std::strtoul(argv[++i], nullptr, 10);
            if (v < 1ul || v > 240ul) {
                std::fprintf(stderr, "hip_cli: --mp4-fps must be in [1, 240]\n");
                return 2;
            }
            mp4_fps = std::uint32_t(v);
        } else if (a == "--precision" && has) {
            const std::string_view p = argv[++i];
            if (p == "fp32") opts.precision = SING_PREC_FP32;
            else if (p == "fp64") opts.precision = SING_PREC_FP64;
            else if (p == "mixed") opts.precision = SING_PREC_MIXED;
            else {
                std::fprintf(stderr, "hip_cli: --precision must be fp32, fp64 or mixed\n");
                return 2;
            }
        } else if (a == "--hdr-output" && has) hdr_path = argv[++i];
        else if (a == "--class-output" && has) cls_path = argv[++i];
        else if (a == "--bloom") bloom = true;
        else if (a == "--stats") print_stats = true;
        else if (a == "--h-step" && has) h_step = std::strtof(argv[++i], nullptr);
        else if (a == "--max-steps" && has) max_steps = std::uint32_t(std::strtoul(argv[++i], nullptr, 10));
        else if (a == "--escape-radius" && has) escape_r = std::strtof(argv[++i], nullptr);
        else if (a == "--peak-temp-scale" && has) temp_scale = std::strtof(argv[++i], nullptr);
        else if (a == "--turb-seed" && has) turb_seed = std::uint32_t(std::strtoul(argv[++i], nullptr, 10));
        else if (a == "--adapt-tol" && has) opts.adapt_tol = std::strtof(argv[++i], nullptr);
        else if (a == "--sampler" && has) {
            const std::string_view p = argv[++i];
            if (p == "halton") opts.sampler = SING_SAMPLER_HALTON;
            else if (p == "grid") opts.sampler = SING_SAMPLER_GRID;
            else {
                std::fprintf(stderr, "hip_cli: --sampler must be halton or grid\n");
                return 2;
            }
        } else if (a == "--refine" && has) {
            std::string_view rest = argv[++i];
            while (!rest.empty()) {
                const std::size_t comma = rest.find(',');
                const std::string_view item = rest.substr(0, comma);
                if (item == "cross64") opts.refine |= SING_REFINE_CROSS64;
                else if (item == "reproject") opts.refine |= SING_REFINE_REPROJECT;
                else {
                    std::fprintf(stderr, "hip_cli: --refine takes cross64 and/or reproject\n");
                    return 2;
                }
                rest = comma == std::string_view::npos ? std::string_view() : rest.substr(comma + 1);
            }
        }
        else {
            std::fprintf(stderr, "hip_cli: unknown argument: %.*s\n", int(a.size()), a.data());
            print_help();
            return 2;
        }
    }

    SingScene scene;
    sing_scene_defaults(&scene);
    if (!scene_path.empty()) {
        SingSceneConfig cfg;
        sing_scene_config_defaults(&cfg);
        int line = 0;
        char err[512] = {0};
        const int rc = sing_load_scene_config(scene_path.c_str(), &cfg, &line, err, sizeof(err));
        if (rc != SING_OK) {
            const char* reason = rc == SING_ERR_IO ? "file-not-found" : rc == SING_ERR_KEY ? "unknown-key" : "parse-error";
            std::fprintf(stderr, "hip_cli: scene load failed (%s, line %d): %s\n", reason, line, err);
            return 2;
        }
        sing_scene_from_config(&cfg, &scene);
    }
    scene.supersample = spp;
    scene.h_step = h_step;
    scene.max_steps = max_steps;
    scene.escape_r_M = escape_r;

    if (frames > 1 && output_pattern.empty()) {
        std::fprintf(stderr, "hip_cli: --frames > 1 requires --output-pattern with a printf-style %%d placeholder, "
                             "e.g. frames/frame_%%04d.png\n");
        return 2;
    }
    if (frames > 1 && output_pattern.find('%') == std::string::npos) {
        std::fprintf(stderr, "hip_cli: --output-pattern must contain a printf placeholder (%%d / %%04d)\n");
        return 2;
    }
    if (!encode_mp4_path.empty() && frames < 2) {
        std::fprintf(stderr, "hip_cli: --encode-mp4 requires --frames > 1 (a sequence to encode)\n");
        return 2;
    }

    SingContext* ctx = sing_create(width, height, 0);
    if (!ctx) {
        std::fprintf(stderr, "hip_cli: backend initialize failed: %s\n", sing_last_error(nullptr));
        return 1;
    }
    opts.outputs = SING_OUT_RGBA8 | (hdr_path.empty() ? 0u : SING_OUT_HDR) | (cls_path.empty() ? 0u : SING_OUT_CLASS);
    opts.post = bloom ? SING_POST_BLOOM : SING_POST_NONE;
    opts.peak_T_scale = temp_scale;
    opts.turb_seed = turb_seed;

    constexpr float kTwoPi = 6.28318530717958647693f;
    float frame_time = 0.0f;
    // Frame f+1 is submitted (sing_render_async) as soon as frame f's outputs
    // are captured, so its trace runs on the GPU while the host writes frame
    // f's files; the files, their order and the printed lines are those of a
    // render-then-write loop (cuda_cli/main.cpp:243-299).
    auto submit = [&](std::uint32_t f) {
        const float az = camera_azimuth + (frames > 1 ? (kTwoPi * float(f) / float(frames)) : 0.0f);
        SingCamera cam;
        sing_orbital_camera(camera_distance, camera_elevation, az, fov_y_deg, &cam);
        SingUniforms u;
        sing_pack_uniforms(&scene, &cam, width, height, frame_time, &u);
        frame_time += 1.0f / 30.0f;
        return sing_render_async(ctx, &u, &opts);
    };
    auto fail_frame = [&](std::uint32_t f) {
        std::fprintf(stderr, "hip_cli: render failed (frame %u): %s\n", f, sing_last_error(ctx));
        sing_destroy(ctx);
        return 1;
    };
    std::vector<std::uint8_t> rgba(std::size_t(width) * height * 4);
    std::vector<float> hdr(hdr_path.empty() ? 0 : std::size_t(width) * height * 3);
    std::vector<std::uint8_t> cls(cls_path.empty() ? 0 : std::size_t(width) * height);
    if (submit(0) != SING_OK) return fail_frame(0);
    for (std::uint32_t f = 0; f < frames; ++f) {
        SingRenderStats st{};
        if (sing_wait(ctx, &st) != SING_OK || sing_capture_rgba8(ctx, rgba.data(), rgba.size()) != SING_OK)
            return fail_frame(f);
        const bool hdr_ok = hdr_path.empty() || sing_capture_hdr(ctx, hdr.data(), hdr.size() * 4) == SING_OK;
        const bool cls_ok = cls_path.empty() || sing_capture_class(ctx, cls.data(), cls.size()) == SING_OK;
        // the next frame renders while this one's files are written; a
        // submission error is reported after them, where a serial loop would
        const bool next_ok = f + 1 >= frames || submit(f + 1) == SING_OK;
        char path_buf[1024];
        const char* path;
        if (frames > 1) {
            std::snprintf(path_buf, sizeof(path_buf), output_pattern.c_str(), int(f));
            path = path_buf;
        } else {
            path = output_path.c_str();
        }
        if (!write_png(path, rgba.data(), width, height)) {
            std::fprintf(stderr, "hip_cli: failed to write '%s'\n", path);
            sing_destroy(ctx);
            return 1;
        }
        if (!hdr_path.empty() && (!hdr_ok || !write_pfm(hdr_path.c_str(), hdr.data(), width, height))) {
            std::fprintf(stderr, "hip_cli: failed to write '%s'\n", hdr_path.c_str());
            sing_destroy(ctx);
            return 1;
        }
        if (!cls_path.empty() && (!cls_ok || !write_pgm(cls_path.c_str(), cls.data(), width, height))) {
            std::fprintf(stderr, "hip_cli: failed to write '%s'\n", cls_path.c_str());
            sing_destroy(ctx);
            return 1;
        }
        if (frames > 1) {
            std::printf("wrote %s (%ux%u, frame %u/%u, HIP)\n", path, width, height, f + 1, frames);
            std::fflush(stdout);
        } else {
            std::printf("wrote %s (%ux%u, HIP)\n", path, width, height);
        }
        if (print_stats)
            std::printf("stats: frame %u kernel %.2f ms trace %.2f ms rays %llu steps %llu fp64 %llu (%.1f Mrays/s)\n", f,
                        st.kernel_ms, st.trace_ms, (unsigned long long)st.rays, (unsigned long long)st.steps,
                        (unsigned long long)st.fp64_rays, double(st.rays) / st.kernel_ms / 1e3);
        if (!next_ok) return fail_frame(f + 1);
    }
    sing_destroy(ctx);

    if (!encode_mp4_path.empty()) {
        char cmd[2048];
        const int n = std::snprintf(cmd, sizeof(cmd),
                                    "ffmpeg -y -hide_banner -loglevel warning -framerate %u -i \"%s\" -c:v libx264 "
                                    "-pix_fmt yuv420p -crf 18 \"%s\"",
                                    mp4_fps, output_pattern.c_str(), encode_mp4_path.c_str());
        if (n < 0 || n >= int(sizeof(cmd))) {
            std::fprintf(stderr, "hip_cli: ffmpeg command exceeded buffer (paths too long)\n");
            return 1;
        }
        std::printf("encoding: %s\n", cmd);
        std::fflush(stdout);
        const int rc = std::system(cmd);
        if (rc != 0) {
            std::fprintf(stderr, "hip_cli: ffmpeg exited %d -- is `ffmpeg` on PATH? (see manual one-liner in --help)\n",
                         rc);
            return 1;
        }
        std::printf("wrote %s (%u frames at %u fps, libx264 crf=18 yuv420p)\n", encode_mp4_path.c_str(), frames,
                    mp4_fps);
    }
    return 0;
}
