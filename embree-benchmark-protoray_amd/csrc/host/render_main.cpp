// render_main.cpp -- `protoray render <scene.mesh> -dev hip -r diffuseHip ...`
//
// The offscreen path of the reference's render command: option handling of
// main/render/render.cpp:49-134,138-240, then the frame loop, warm-up and
// statistics of RenderWindow (main/render/render_window.cpp:28-162,164-305,
// 307-325,346-374,400-426) and StatsRecorder (main/render/stats_recorder.cpp).
// Benchmark mode writes <prefix>.txt (stats_<key>=<value>), <prefix>.csv,
// <prefix>.ppm and <prefix>.log, so benchmark/bench-cpugpu.sh-style scripts
// read the results unchanged.  Only "-dev hip" exists in this build.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "device_hip.h"
#include "mesh_build.h"
#include "mesh_io.h"

using namespace prt;

namespace {

struct Option {
    std::string name, value;
};

// -name value | -name=value | positional (sys/option.cpp:37-93)
std::vector<Option> parse_options(const std::vector<std::string>& args)
{
    std::vector<Option> opts;
    bool pending = false;
    Option opt;
    for (const std::string& a : args) {
        if (a.size() > 1 && a[0] == '-' && std::isalpha((unsigned char)a[1])) {
            if (pending) opts.push_back(opt);
            size_t s = a.find_first_not_of('-');
            std::string body = a.substr(s);
            size_t eq = body.find('=');
            if (eq != std::string::npos) {
                opts.push_back({body.substr(0, eq), body.substr(eq + 1)});
                pending = false;
            } else {
                opt = {body, ""};
                pending = true;
            }
        } else {
            if (!pending) opt.name.clear();
            opt.value = a;
            opts.push_back(opt);
            pending = false;
        }
    }
    if (pending) opts.push_back(opt);
    return opts;
}

std::vector<std::string> read_config(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("could not open config: " + path);
    std::vector<std::string> out;
    std::string tok;
    while (f >> tok) out.push_back(tok);
    return out;
}

std::ofstream g_log;

void log_line(const std::string& s)
{
    std::cout << s << std::endl;
    if (g_log) g_log << s << std::endl;
}

std::string fmt(double v)
{
    std::ostringstream sm;
    sm.precision(10);
    sm << v;
    return sm.str();
}

struct Settings {
    std::string scene, device = "hip", renderer = "diffuseHip", prefix = "render", input_path;
    Vec2i size{1024, 576};
    int view_id = 0;
    bool benchmark = false;
    Props props;
};

bool apply_options(const std::vector<Option>& opts, Settings& st)
{
    for (const Option& o : opts) {
        if (o.name.empty()) st.scene = st.input_path + o.value;
        else if (o.name == "c" || o.name == "config") {
            size_t slash = o.value.find_last_of('/');
            st.input_path = slash == std::string::npos ? "" : o.value.substr(0, slash + 1);
            if (!apply_options(parse_options(read_config(o.value)), st)) return false;
        } else if (o.name == "dev" || o.name == "device" || o.name == "e" || o.name == "engine") st.device = o.value;
        else if (o.name == "r" || o.name == "renderer") st.renderer = o.value;
        else if (o.name == "a" || o.name == "accel") st.props.set("accel", o.value);
        else if (o.name == "s" || o.name == "size") st.size = Value(o.value).get<Vec2i>();
        else if (o.name == "benchmark" || o.name == "bench") {
            st.benchmark = true;
            st.prefix = o.value.empty() ? "benchmark" : o.value;
            st.props.set("benchmark", st.prefix);
        } else if (o.name == "o" || o.name == "out" || o.name == "output") st.prefix = o.value;
        else if (o.name == "view") {
            st.view_id = Value(o.value).get<int>();
            if (st.view_id < 0 || st.view_id >= kViewSetSize) {
                std::cerr << "Error: Invalid view" << std::endl;
                return false;
            }
        } else if (o.name == "do" || o.name == "dw" || o.name == "df") {
            // display mode: only offscreen rendering exists here
        } else st.props.set(o.name, o.value);
    }
    return true;
}

void write_ppm(const std::string& path, const std::vector<int32_t>& px, int w, int h)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return;
    fprintf(f, "P6\n%d %d\n255\n", w, h);
    for (int i = 0; i < w * h; ++i) {
        const uint32_t p = (uint32_t)px[i];
        fputc(p & 0xff, f);
        fputc((p >> 8) & 0xff, f);
        fputc((p >> 16) & 0xff, f);
    }
    fclose(f);
}

// The camera of view st.view_id: <scene>.view if present, else the default
// view (RenderWindow::resetView, main/render/render_window.cpp:400-426),
// through RenderWindow::makeCamera (main/render/render_window.cpp:307-325).
void init_view(DeviceHip& device, const Settings& st)
{
    const std::string base = st.scene.substr(0, st.scene.find_last_of('.'));
    View views[kViewSetSize];
    const Box3f b = device.getSceneBounds();
    const Vec3f c{0.5f * (b.low.x + b.high.x), 0.5f * (b.low.y + b.high.y), 0.5f * (b.low.z + b.high.z)};
    const float ext = std::fmax(std::fmax(b.high.x - b.low.x, b.high.y - b.low.y), b.high.z - b.low.z);
    View def;
    def.pos = Vec3f{c.x, c.y, 1.7f * ext + c.z};
    def.fov_y = 45.0f * (3.14159265358979323846f / 180.0f);
    def.focus = std::sqrt((c.z - def.pos.z) * (c.z - def.pos.z));
    for (View& v : views) v = def;
    load_view_set(base + ".view", views);
    const View& view = views[st.view_id];

    float basis[9];
    prt_hip_view_basis(view.angle_x, view.angle_y, basis);
    Props cam;
    cam.set("type", view.radius == 0.0f ? "pinhole" : "thinlens");
    cam.set("position", view.pos);
    cam.set("basis", Basis3f{{basis[0], basis[1], basis[2]}, {basis[3], basis[4], basis[5]},
                             {basis[6], basis[7], basis[8]}});
    cam.set("fov", view.fov_y);
    cam.set("aspectRatio", (float)st.size.x / (float)st.size.y);
    cam.set("lensRadius", view.radius);
    cam.set("focalDistance", view.focus);
    device.initCamera(cam);
}

void print_query(int i, const Props& res)
{
    std::cout << "query " << i << std::endl;
    for (const auto& kv : res.items()) std::cout << kv.first << "=" << kv.second.text() << std::endl;
    std::cout << "end" << std::endl;
}

}  // namespace

// `protoray query <scene.mesh> [-size W,H] [-view N] [-pixel x,y]... [-ray ox,oy,oz,dx,dy,dz]...`
// Device::queryPixel / Device::queryRay (render/device.h:43-44), which the
// reference's interactive window calls on a click (RenderWindow, main/render/
// render_window.cpp:447); here for scripts and tests.  Prints, per query in
// order, "query <i>", the result's key=value lines (none for a miss), "end".
int query_main(const std::vector<std::string>& args)
{
    Settings st;
    std::vector<std::pair<bool, std::vector<double>>> queries;  // (is pixel, numbers)
    std::vector<Option> rest;
    try {
        for (const Option& o : parse_options(args)) {
            if (o.name == "pixel" || o.name == "ray") {
                std::vector<double> n;
                std::string v = o.value;
                for (char& ch : v)
                    if (ch == ',') ch = ' ';
                std::istringstream sm(v);
                double d;
                while (sm >> d) n.push_back(d);
                if (n.size() != (o.name == "pixel" ? 2u : 6u)) {
                    std::cerr << "Error: -" << o.name << " needs " << (o.name == "pixel" ? 2 : 6) << " numbers"
                              << std::endl;
                    return 1;
                }
                queries.emplace_back(o.name == "pixel", n);
            } else {
                rest.push_back(o);
            }
        }
        if (!apply_options(rest, st)) return 1;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    if (st.scene.empty()) {
        std::cerr << "Error: Scene not specified" << std::endl;
        return 1;
    }
    try {
        DeviceHip device(0);
        device.initScene(st.scene, st.props);
        device.initFrame(st.size, st.props);
        Props rprops = st.props;
        rprops.set("type", st.renderer);
        rprops.set("imageSize", st.size);
        Props build_stats;
        device.initRenderer(rprops, build_stats);
        init_view(device, st);
        for (size_t i = 0; i < queries.size(); ++i) {
            const std::vector<double>& n = queries[i].second;
            if (queries[i].first) {
                print_query((int)i, device.queryPixel((int)n[0], (int)n[1]));
            } else {
                Ray r;
                r.org = Vec3f{(float)n[0], (float)n[1], (float)n[2]};
                r.dir = Vec3f{(float)n[3], (float)n[4], (float)n[5]};
                print_query((int)i, device.queryRay(r));
            }
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

int render_main(const std::vector<std::string>& args)
{
    std::cout << "ProtoRay Render (HIP)" << std::endl << std::endl;
    if (args.empty()) {
        std::cout << "Usage: protoray render [options]" << std::endl;
        return 0;
    }
    Settings st;
    try {
        if (!apply_options(parse_options(args), st)) return 1;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    if (st.scene.empty()) {
        std::cerr << "Error: Scene not specified" << std::endl;
        return 1;
    }
    if (st.size.x <= 0 || st.size.x % 8 != 0 || st.size.y <= 0 || st.size.y % 8 != 0) {
        std::cerr << "Error: Resolution must be multiple of 8" << std::endl;
        return 1;
    }
    if (st.device != "hip") {
        std::cerr << "Error: Invalid device: " << st.device << std::endl;
        return 1;
    }
    g_log.open(st.benchmark ? st.prefix + ".log" : std::string("render.log"));

    try {
        DeviceHip device(0);
        log_line("Device: " + device.getInfo());
        device.initScene(st.scene, st.props);
        device.initFrame(st.size, st.props);
        Props rprops = st.props;
        rprops.set("type", st.renderer);
        rprops.set("imageSize", st.size);
        Props build_stats;
        device.initRenderer(rprops, build_stats);
        log_line("Build time: " + fmt(build_stats.get<double>("buildMs")) + " ms");

        init_view(device, st);
        device.clearFrame();

        const int max_spp = st.props.get("spp", 1);
        const int warmup = st.props.get("warmup", max_spp / 6);
        const bool checkpoint = st.props.exists("checkpoint");
        std::vector<std::string> names;
        std::map<std::string, std::vector<double>> recorded;
        std::vector<int32_t> px((size_t)st.size.x * st.size.y);
        auto save_image = [&](const std::string& path) {
            Surface s{px.data(), st.size.x, st.size.y, st.size.x * 4};
            device.blitFrame(s);
            write_ppm(path, px, st.size.x, st.size.y);
        };
        for (int spp = 0; spp < max_spp;) {
            Props fs;
            const auto t0 = std::chrono::steady_clock::now();
            device.render(fs);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            fs.set("renderMs", ms);
            fs.set("fps", 1000.0 / ms);
            ++spp;
            if (spp > warmup || !st.benchmark) {
                for (const auto& kv : fs.items()) {
                    if (!recorded.count(kv.first)) names.push_back(kv.first);
                    recorded[kv.first].push_back(kv.second.get<double>());
                }
            }
            if (checkpoint && (spp & (spp - 1)) == 0) {
                char name[64];
                snprintf(name, sizeof(name), "%06dspp", spp);
                save_image(st.prefix + "_" + name + ".ppm");
            }
        }

        Props avg = build_stats;
        for (const std::string& n : names) {
            double s = 0.0;
            for (double v : recorded[n]) s += v;
            avg.set(n, s / (double)recorded[n].size());
        }
        std::string line = "Average:";
        for (const auto& kv : avg.items()) line += " " + kv.first + "=" + fmt(kv.second.get<double>());
        log_line(line);
        if (st.benchmark) {
            FILE* f = fopen((st.prefix + ".txt").c_str(), "wt");
            if (f) {
                for (const auto& kv : avg.items())
                    fprintf(f, "stats_%s=%s\n", kv.first.c_str(), fmt(kv.second.get<double>()).c_str());
                fclose(f);
            }
            std::ofstream csv(st.prefix + ".csv");
            for (size_t i = 0; i < names.size(); ++i) csv << (i ? "," : "") << names[i];
            csv << "\n";
            size_t rows = names.empty() ? 0 : recorded[names[0]].size();
            for (size_t r = 0; r < rows; ++r) {
                for (size_t i = 0; i < names.size(); ++i) csv << (i ? "," : "") << fmt(recorded[names[i]][r]);
                csv << "\n";
            }
        }
        save_image(st.prefix + ".ppm");
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// `protoray build <input.obj|input.mesh> [-o output.mesh] [-t mesh]`
// (mainBuild, main/build/build.cpp:31-130): OBJ through the loader and the
// mesh builder, a .mesh loaded as is; the result saved as a .mesh blob.
int build_main(const std::vector<std::string>& args)
{
    std::cout << "ProtoRay Build" << std::endl << std::endl;
    if (args.empty()) {
        std::cout << "Usage: protoray build [options]" << std::endl;
        return 0;
    }
    std::string type = "mesh", input, output;
    for (const Option& o : parse_options(args)) {
        if (o.name.empty()) {
            if (!input.empty()) {
                std::cerr << "Error: Multiple inputs" << std::endl;
                return 1;
            }
            input = o.value;
        } else if (o.name == "o") {
            if (!output.empty()) {
                std::cerr << "Error: Multiple outputs" << std::endl;
                return 1;
            }
            output = o.value;
        } else if (o.name == "t") {
            type = o.value;
        } else {
            std::cout << "Option: " << o.name << " = " << o.value << std::endl;
        }
    }
    const size_t dot = input.find_last_of('.');
    const size_t slash = input.find_last_of('/');
    const bool has_ext = dot != std::string::npos && (slash == std::string::npos || dot > slash);
    const std::string ext = has_ext ? input.substr(dot + 1) : "";
    if (output.empty()) output = (has_ext ? input.substr(0, dot) : input) + "." + type;
    try {
        TriangleMesh mesh;
        if (ext == "mesh") {
            mesh = load_mesh(input);
        } else if (ext == "obj") {
            int degenerate = 0;
            const TriangleSoup soup = load_obj(input, &degenerate);
            if (degenerate > 0) std::cout << "Warning: Removed " << degenerate << " degenerate triangles" << std::endl;
            std::cout << "Building mesh: " << soup.triangles.size() << " triangles, " << soup.triangles.size() * 3
                      << " vertices" << std::endl;
            mesh = build_mesh(soup);
            std::cout << "Unique vertices: " << mesh.vertex_count() << std::endl;
        } else {
            std::cerr << "Error: Unsupported input format" << std::endl;
            return 1;
        }
        if (type != "mesh") {
            std::cerr << "Error: Invalid type" << std::endl;
            return 1;
        }
        save_mesh(output, mesh);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

int main(int argc, char* argv[])
{
    const std::string cmd = argc < 2 ? "" : argv[1];
    std::vector<std::string> args(argv + (argc < 2 ? argc : 2), argv + argc);
    if (cmd == "render") return render_main(args);
    if (cmd == "build") return build_main(args);
    if (cmd == "query") return query_main(args);
    std::cout << "Usage: protoray render <scene.mesh> [-dev hip] [-r primaryHip|diffuseHip|diffuse2Hip] [-size W,H] "
                 "[-spp N] [-maxDepth D] [-albedo R] [-view N] [-benchmark prefix]" << std::endl
              << "       protoray build <scene.obj|scene.mesh> [-o out.mesh] [-t mesh]" << std::endl
              << "       protoray query <scene.mesh> [-size W,H] [-view N] [-pixel x,y]... "
                 "[-ray ox,oy,oz,dx,dy,dz]..." << std::endl;
    return argc < 2 ? 0 : 1;
}
