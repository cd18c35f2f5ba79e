// gammaray -- the reference binary's progressive render loop
// (src/bin/gammaray.rs:7-113) over the C ABI: the demo scene (the dragon OBJ,
// a light sphere, a large metal sphere, a glass sphere and two rows of 20
// spheres), Camera::default, 704x512, one BDPT iteration per pass, the film
// rewritten as output.exr after every pass and a timing line per pass.
//
// The film's host copy is overlapped with rendering: after pass k is queued,
// its snapshot is copied to host memory on a side stream while pass k+1
// renders, and a writer thread encodes and writes the EXR of pass k.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gammaray.h"

namespace {

constexpr float kHorizontalAperture35mm = 2.2f;  // src/core/camera.rs:8-9
constexpr float kVerticalAperture35mm = 1.6f;

struct Disney {  // DisneyMaterialBuilder defaults, src/material/material.rs:276-290
    float base[3] = {1.0f, 1.0f, 1.0f};
    float roughness = 0.5f, anisotropic = 0.0f, ior = 1.5f, metallic = 0.0f, specular_trans = 0.0f;
    float specular_tint = 0.0f, sheen = 0.0f, sheen_tint = 0.5f, clearcoat = 0.0f, clearcoat_gloss = 0.1f;
    void record(float* r) const {
        std::memset(r, 0, GR_MAT_FLOATS * sizeof(float));
        r[0] = 0.0f;
        r[1] = base[0], r[2] = base[1], r[3] = base[2];
        r[4] = roughness, r[5] = anisotropic, r[6] = ior, r[7] = metallic, r[8] = specular_trans;
        r[9] = specular_tint, r[10] = sheen, r[11] = sheen_tint, r[12] = clearcoat, r[13] = clearcoat_gloss;
    }
};

int check(int code, gr_ctx* ctx, const char* what) {
    if (code != GR_OK) {
        const char* msg = ctx ? gr_ctx_last_error(ctx) : nullptr;
        std::fprintf(stderr, "gammaray: %s failed (%d)%s%s\n", what, code, msg ? ": " : "", msg ? msg : "");
        std::exit(1);
    }
    return code;
}

void sphere(gr_scene* s, gr_ctx* ctx, float x, float y, float z, float r, const float* mat) {
    const float c[3] = {x, y, z};
    check(gr_scene_add_sphere(s, c, r, mat, nullptr), ctx, "add sphere");
}

// The scene of src/bin/gammaray.rs:8-81.
void build_demo(gr_scene* s, gr_ctx* ctx, const std::string& obj, bool obj_required) {
    float mat[GR_MAT_FLOATS];
    {
        Disney d;
        d.base[0] = 0.0f, d.base[1] = 0.5f, d.base[2] = 1.0f;
        d.roughness = 0.5f, d.metallic = 1.0f, d.anisotropic = 1.0f, d.sheen = 1.0f;
        d.record(mat);
        // Mat::scale(0.9) * Mat::translation(-4, -5, -100), row-major
        const float xf[16] = {0.9f, 0, 0, 0, 0, 0.9f, 0, 0, 0, 0, 0.9f, 0, -4.0f, -5.0f, -100.0f, 1.0f};
        gr_obj* o = nullptr;
        char err[512] = {0};
        if (gr_obj_load(obj.c_str(), xf, &o, err, sizeof(err)) == GR_OK) {
            check(gr_scene_add_obj(s, o, mat, nullptr), ctx, "add mesh");
            gr_obj_destroy(o);
        } else if (obj_required) {
            std::fprintf(stderr, "gammaray: %s\n", err);
            std::exit(1);
        } else {
            std::fprintf(stderr, "gammaray: %s -- rendering the scene without the mesh\n", err);
        }
    }
    float light[GR_MAT_FLOATS] = {1.0f, 2.0f, 2.0f, 2.0f};  // Material::diffuse_light
    sphere(s, ctx, 12.0f, 3.0f, -90.0f, 5.0f, light);
    {
        Disney d;
        d.base[0] = 0.5f, d.base[1] = 0.9f, d.base[2] = 0.0f;
        d.roughness = 0.5f, d.metallic = 1.0f;
        d.record(mat);
        sphere(s, ctx, -25.0f, 0.0f, -50.0f, 75.0f, mat);
    }
    {
        Disney d;
        d.specular_trans = 1.0f, d.roughness = 0.2f, d.ior = 1.8f, d.metallic = 0.0f;
        d.record(mat);
        sphere(s, ctx, 6.0f, -10.0f, -90.0f, 4.0f, mat);
    }
    for (int row = 0; row < 2; ++row) {
        for (int i = 0; i < 20; ++i) {
            Disney d;
            d.base[0] = i % 2 == 0 ? 0.9f : 0.2f;
            d.base[1] = i % 3 == 0 ? 0.9f : 0.2f;
            d.base[2] = i % 5 == 0 ? 0.9f : 0.2f;
            if (row == 0) {
                d.roughness = 0.5f, d.metallic = 1.0f, d.anisotropic = 0.8f, d.clearcoat = 1.0f, d.clearcoat_gloss = 1.0f;
            } else {
                d.roughness = 0.2f, d.ior = 1.8f, d.metallic = 0.0f, d.sheen = 1.0f;
            }
            d.record(mat);
            sphere(s, ctx, -20.0f + (float)(i * 4), row == 0 ? 12.0f : -12.0f, -100.0f, 1.8f, mat);
        }
    }
}

// Double-buffered EXR writer: the render loop hands over a host film copy and
// goes on; the writer encodes and rewrites the file (ExrWriter::update + write).
class Writer {
public:
    Writer(std::string path, uint32_t w, uint32_t h) : path_(std::move(path)), w_(w), h_(h), thread_([this] { run(); }) {}
    ~Writer() { close(); }
    // Writes what is still pending, then stops the thread.
    void close() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
    }
    // Takes ownership of px (swapped out); waits while the previous copy is still being written.
    void submit(std::vector<float>& px) {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [this] { return !full_; });
        pending_.swap(px);
        full_ = true;
        cv_.notify_all();
    }
    bool ok() const { return ok_; }  // after close()

private:
    void run() {
        std::vector<float> px;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return full_ || stop_; });
                if (!full_) return;
                px.swap(pending_);
                full_ = false;
                cv_.notify_all();
            }
            if (gr_exr_write(px.data(), w_, h_, path_.c_str()) != GR_OK) {
                std::fprintf(stderr, "gammaray: cannot write %s\n", path_.c_str());
                ok_ = false;
            }
        }
    }
    std::string path_;
    uint32_t w_, h_;
    std::mutex m_;
    std::condition_variable cv_;
    std::vector<float> pending_;
    bool full_ = false, stop_ = false;
    bool ok_ = true;
    std::thread thread_;
};

void usage() {
    std::fprintf(stderr,
                 "usage: gammaray [--iterations N] [--height H] [--integrator bdpt|bdpt_mis|bdpt_lt|pt|pt_nee|display]\n"
                 "                [--obj PATH] [--output PATH] [--seed S] [--max-depth D] [--light-depth D]\n"
                 "                [--device N]\n");
}

}  // namespace

int main(int argc, char** argv) {
    uint32_t limit = 200, height = 512, max_depth = 0, light_depth = 0;
    int device = 0;
    uint64_t seed = 1;
    std::string obj = "dragon100k_uvs.obj", output = "output.exr", integ = "bdpt";
    bool obj_required = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--iterations") limit = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--height") height = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--integrator") integ = next();
        else if (a == "--obj") obj = next(), obj_required = true;
        else if (a == "--output") output = next();
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--max-depth") max_depth = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--light-depth") light_depth = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--device") device = std::atoi(next());
        else {
            usage();
            return a == "--help" || a == "-h" ? 0 : 2;
        }
    }
    uint32_t integrator;
    if (integ == "bdpt") integrator = GR_INTEGRATOR_BDPT;
    else if (integ == "pt") integrator = GR_INTEGRATOR_PT;
    else if (integ == "pt_nee") integrator = GR_INTEGRATOR_PT_NEE;
    else if (integ == "display") integrator = GR_INTEGRATOR_DISPLAY;
    else if (integ == "bdpt_mis") integrator = GR_INTEGRATOR_BDPT_MIS;
    else if (integ == "bdpt_lt") integrator = GR_INTEGRATOR_BDPT_LT;
    else {
        usage();
        return 2;
    }
    if (height == 0) {
        usage();
        return 2;
    }

    gr_camera cam;  // Camera::default (src/core/camera.rs:37-40)
    std::memset(&cam, 0, sizeof(cam));
    cam.focal_length = 5.0f;
    cam.horizontal_aperture = kHorizontalAperture35mm;
    cam.vertical_aperture = kVerticalAperture35mm;
    cam.f_stop = 8.0f;
    cam.rotate[0] = 1.0f;
    const float aspect = cam.horizontal_aperture / cam.vertical_aperture;
    const uint32_t width = (uint32_t)((float)height * aspect);
    std::printf("Aspect ratio: %g, Width: %u, Height: %u\n", aspect, width, height);

    gr_ctx* ctx = nullptr;
    check(gr_ctx_create(device, &ctx), nullptr, "context");
    gr_scene* scene = nullptr;
    check(gr_scene_create(ctx, &scene), ctx, "scene");
    build_demo(scene, ctx, obj, obj_required);
    check(gr_scene_commit(scene), ctx, "scene commit");
    gr_film* film = nullptr;
    check(gr_film_create(ctx, width, height, &film), ctx, "film");

    gr_trace_params prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.integrator = integrator;
    prm.spp = 1;
    prm.max_depth = max_depth;
    prm.light_depth = light_depth;
    prm.seed = seed;
    std::vector<float> px((size_t)width * height * 4);
    int rc = 0;
    {
        Writer writer(output, width, height);
        double total = 0.0;
        auto last = std::chrono::steady_clock::now();
        for (uint32_t it = 0; it < limit; ++it) {
            // pass it is queued; the readback of pass it-1 finishes meanwhile
            prm.sample_offset = it;
            check(gr_trace(scene, &cam, &prm, film), ctx, "trace");
            if (it > 0) {
                check(gr_film_readback_end(film, px.data()), ctx, "readback");
                writer.submit(px);
                px.resize((size_t)width * height * 4);
            }
            check(gr_film_readback_begin(film), ctx, "readback");
            if (it > 0) {
                const auto now = std::chrono::steady_clock::now();
                const double secs = std::chrono::duration<double>(now - last).count();
                last = now;
                total += secs;
                std::printf("Iteration %u [duration: %.3f sec / %.3f fps] [total: %.3f sec]\n", it - 1, secs,
                            1.0 / secs, total);
                std::fflush(stdout);
            }
        }
        if (limit > 0) {
            check(gr_film_readback_end(film, px.data()), ctx, "readback");
            writer.submit(px);
            const auto now = std::chrono::steady_clock::now();
            const double secs = std::chrono::duration<double>(now - last).count();
            total += secs;
            std::printf("Iteration %u [duration: %.3f sec / %.3f fps] [total: %.3f sec]\n", limit - 1, secs,
                        1.0 / secs, total);
        }
        writer.close();
        rc = writer.ok() ? 0 : 1;
    }
    gr_film_destroy(film);
    gr_scene_destroy(scene);
    gr_ctx_destroy(ctx);
    return rc;
}
