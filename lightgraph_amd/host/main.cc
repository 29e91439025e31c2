// Command line of the LightGraph renderer (main.cc of the reference): the same
// options and help aliases, the stages in the same order with the same stage
// timing lines, on the MI355X path.  -vdbIn reads the density and temperature
// FloatGrids of a .vdb file (host/vdb.cpp) into dense bricks.  Additions:
// -light point:..., procedural scenes (-scene, -sceneSize, -sceneSeed), graph
// size and radius (-points, -radius, -maxConnect) and brick precision (-precision).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>

#include "lightgraph.hpp"
#include "vdb.hpp"

using namespace lightgraph;
using std::cout;
using std::endl;
using std::string;

namespace {

void printHelp() {
    const char *lines[][3] = {
        {"-vdbIn", "Input vdb file to read and render", "string"},
        {"-exrOut", "Output exr file to write to disk", "string"},
        {"-camPos", "Camera position", "float,float,float"},
        {"-camRot", "Camera rotation", "float,float,float"},
        {"-camYRot", "World centered camera Y rotation (turntable)", "float"},
        {"-resWidth", "Output image resolution in width", "int"},
        {"-resHeight", "Output image resolution in height", "int"},
        {"-focal", "Camera focal length in mm (full frame sensor)", "float"},
        {"-samples", "Number of antialising samples per pixel", "int"},
        {"-primaryStep", "Number of voxel width per primary sample in the volume", "float"},
        {"-shadowStep", "Number of voxel width per shadow sample in the volume", "float"},
        {"-cutoff", "Density threshold, values lower are ignored", "float"},
        {"-light", "Add a light to the scene",
         "dir:yRot:r,g,b:intensity:rx,ry,rz | point:yRot:r,g,b:intensity:x,y,z | "
         "env:yRot:r,g,b:intensity:samples,/path/to/envMap.exr"},
        {"-iter", "Number of LightGraph iteration", "int"},
        {"-scatterGridResFactor", "Factor by which the scatter grid is scaled down", "int"},
        {"-ptsPerKernel", "Number of points used per kernel for a single iteration", "int"},
        {"-densityScale", "Density multiplier", "float"},
        {"-densityMax", "Make density values lower or equal to this value", "float"},
        {"-shadowDensityScale", "Density multiplier for the shadows only", "float"},
        {"-scatterDensityScale", "Density multiplier for the scatter only", "float"},
        {"-scatterDensityMin", "Make density values greater or equal to this value for scatter only", "float"},
        {"-absorption", "Shader absorption parameter", "float,float,float"},
        {"-scattering", "Shader scattering parameter", "float,float,float"},
        {"-volumeColor", "Shader volume color", "float,float,float"},
        {"-emissionScale", "Shader emission multiplier", "float"},
        {"-scatterScale", "Shader scattering multiplier", "float"},
        {"-scatterDensityMaskPower", "Power applied to the density field before masking the scatter field", "float"},
        {"-tempColorRamp", "Define a color ramp to remap temperature values to emission colors",
         "float,...,float/float,float,float:...:float,float,float/float/float"},
        {"-v", "Enable verbose output", ""},
        {"-geoDump", "Write LightGraph geometry to disk as ASCII files", ""},
        {"-h", "Print this help", ""},
        {"-help", "Print this help", ""},
        {"--help", "Print this help", ""},
        {"-scene", "Procedural volume instead of a vdb file: sphere | blobs | plume", "string"},
        {"-sceneSize", "Procedural volume side in voxels", "int"},
        {"-sceneSeed", "Procedural volume seed", "int"},
        {"-points", "Graph points per LightGraph iteration", "int"},
        {"-radius", "Graph edge radius in world units (0: from the volume bbox)", "float"},
        {"-maxConnect", "Maximum edges per graph point (<= 0: uncapped)", "int"},
        {"-sampler", "Graph points: density (density-weighted) or dart (the reference's dart thrower)", "string"},
        {"-scatterCount", "Dart throws per iteration with -sampler dart", "int"},
        {"-precision", "Density brick storage: fp32 | fp16 | bf16", "string"},
        {"-compression", "EXR compression: zip | zips | none", "string"},
    };
    for (auto &l : lines) {
        printf("%-26s: %s\n", l[0], l[1]);
        if (l[2][0]) printf("%-26s  format: %s\n", "", l[2]);
    }
}

// strToVec (main.cc:213-229): one value broadcasts.
Vec3d strToVec(const string &s) {
    std::vector<double> v;
    std::stringstream ss(s);
    string item;
    while (std::getline(ss, item, ',')) v.push_back(std::stod(item));
    if (v.size() == 1) return Vec3d(v[0]);
    if (v.size() < 3) throw Error("expected float,float,float: " + s, LG_ERR_INVALID);
    return Vec3d(v[0], v[1], v[2]);
}

double secondsSince(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char *argv[]) {
    SceneSettings opts;
    Shader shader;
    string vdbFilename, imgFilename, scene, precision = "fp32";
    int sceneSize = 128;
    uint64_t sceneSeed = 1;
    std::vector<string> lightSpecs;
    try {
        for (int i = 1; i < argc; ++i) {
            const string arg = argv[i];
            auto next = [&]() -> string {
                if (i + 1 >= argc) throw Error(arg + " needs a value", LG_ERR_INVALID);
                return argv[++i];
            };
            if (arg.empty() || arg[0] != '-') continue;
            if (arg == "-vdbIn") vdbFilename = next();
            else if (arg == "-exrOut") imgFilename = next();
            else if (arg == "-camPos") opts.camTranslate = strToVec(next());
            else if (arg == "-camRot") opts.camRotate = strToVec(next());
            else if (arg == "-camYRot") opts.camYRotate = std::stod(next());
            else if (arg == "-resWidth") opts.width = (uint64_t)std::atoi(next().c_str());
            else if (arg == "-resHeight") opts.height = (uint64_t)std::atoi(next().c_str());
            else if (arg == "-focal") opts.focal = (float)std::atof(next().c_str());
            else if (arg == "-samples") opts.samples = (uint64_t)std::max(0, std::atoi(next().c_str()));
            else if (arg == "-primaryStep") opts.step[0] = std::atof(next().c_str());
            else if (arg == "-shadowStep") opts.step[1] = std::atof(next().c_str());
            else if (arg == "-cutoff") opts.cutoff = std::atof(next().c_str());
            else if (arg == "-light") lightSpecs.push_back(next());
            else if (arg == "-iter") opts.iter = (size_t)std::atoi(next().c_str());
            else if (arg == "-scatterGridResFactor") opts.scatterGridResFactor = (size_t)std::atoi(next().c_str());
            else if (arg == "-ptsPerKernel") opts.ptsPerKernel = (size_t)std::atoi(next().c_str());
            else if (arg == "-densityScale") shader.setDensityScale(std::stod(next()));
            else if (arg == "-densityMax") shader.setDensityMax(std::stod(next()));
            else if (arg == "-shadowDensityScale") shader.setShadowDensityScale(std::stod(next()));
            else if (arg == "-scatterDensityScale") shader.setScatterDensityScale(std::stod(next()));
            else if (arg == "-scatterDensityMin") shader.setScatterDensityMin(std::stod(next()));
            else if (arg == "-absorption") shader.setAbsorption(strToVec(next()));
            else if (arg == "-scattering") shader.setScattering(strToVec(next()));
            else if (arg == "-volumeColor") shader.setVolumeColor(strToVec(next()));
            else if (arg == "-emissionScale") shader.setEmissionScale(std::stod(next()));
            else if (arg == "-scatterScale") shader.setScatterScale(std::stod(next()));
            else if (arg == "-scatterDensityMaskPower") shader.setScatterDensityMaskPower(std::stod(next()));
            else if (arg == "-tempColorRamp") shader.setEmissionColorRamp(next());
            else if (arg == "-v") opts.verbose = true;
            else if (arg == "-geoDump") opts.geoDump = true;
            else if (arg == "-h" || arg == "-help" || arg == "--help") {
                printHelp();
                return 0;
            } else if (arg == "-scene") scene = next();
            else if (arg == "-sceneSize") sceneSize = std::atoi(next().c_str());
            else if (arg == "-sceneSeed") sceneSeed = (uint64_t)std::atoll(next().c_str());
            else if (arg == "-points") opts.points = (uint32_t)std::atoi(next().c_str());
            else if (arg == "-radius") opts.radius = std::atof(next().c_str());
            else if (arg == "-maxConnect") opts.maxConnect = std::atoi(next().c_str());
            else if (arg == "-sampler") opts.sampler = next();
            else if (arg == "-scatterCount") opts.scatterCount = (uint32_t)std::atoi(next().c_str());
            else if (arg == "-precision") precision = next();
            else if (arg == "-compression") opts.compression = next();
            else {
                cout << arg << " is not a valid option" << endl;
                return 1;
            }
        }
        if (scene.empty() && vdbFilename.empty()) {
            cout << "no input: give -vdbIn file.vdb or -scene sphere|blobs|plume" << endl;
            return 1;
        }
        if (!scene.empty() && !vdbFilename.empty()) {
            cout << "give either -vdbIn or -scene, not both" << endl;
            return 1;
        }
        // the dumps' and default image's stem: the vdb file's stem, or the scene's name
        string stem = scene;
        if (!vdbFilename.empty()) {
            stem = vdbFilename.substr(vdbFilename.find_last_of('/') + 1);
            if (stem.size() > 4 && stem.compare(stem.size() - 4, 4, ".vdb") == 0) stem.resize(stem.size() - 4);
        }
        if (imgFilename.empty()) imgFilename = stem + ".exr";
        if (opts.sampler != "density" && opts.sampler != "dart") {
            cout << "-sampler: " << opts.sampler << " is not density or dart" << endl;
            return 1;
        }

        lg_scene_spec spec{};
        spec.n = sceneSize;
        spec.seed = sceneSeed;
        spec.octaves = 6;
        spec.dtype = precision == "fp16" ? LG_F16 : precision == "bf16" ? LG_BF16 : LG_F32;
        if (precision != "fp32" && precision != "fp16" && precision != "bf16")
            throw Error("-precision must be fp32, fp16 or bf16", LG_ERR_INVALID);
        if (scene == "sphere") {
            spec.kind = LG_SCENE_SPHERE_FBM;
            spec.octaves = 4;
            spec.radius = 0.375 * sceneSize;
        } else if (scene == "blobs") {
            spec.kind = LG_SCENE_BLOBS_FBM;
            spec.blobs = 20;
            spec.axis_min = sceneSize * 30.0 / 512.0;
            spec.axis_max = sceneSize * 120.0 / 512.0;
            spec.active_fraction = 0.30;
        } else if (scene == "plume") {
            spec.kind = LG_SCENE_PLUME;
            spec.with_temperature = 1;
        } else if (!scene.empty()) {
            throw Error("-scene must be sphere, blobs or plume", LG_ERR_INVALID);
        }
        opts.geoName = stem;
        // LOAD INPUT GRIDS (main.cc:484-502): the file is read before the device is set up
        VdbVolume vdb;
        if (!vdbFilename.empty()) {
            cout << "Reading in: " << vdbFilename << endl;
            vdb = readVdbVolume(vdbFilename);
        }
        Context ctx(0);
        // lights: dir/point parse on the host; env needs its map on the device
        std::vector<std::unique_ptr<EnvMap>> envMaps;
        for (const string &spec : lightSpecs) {
            lg_light l{};
            if (spec.compare(0, 4, "env:") == 0) {
                // env:yRot:color:intensity:samples,path (main.cc:244-253)
                std::vector<string> f;
                std::stringstream ss(spec);
                string item;
                while (std::getline(ss, item, ':')) f.push_back(item);
                if (f.size() < 5) throw Error("bad env light: " + spec, LG_ERR_INVALID);
                const size_t comma = f[4].find(',');
                if (comma == string::npos) throw Error("env light needs samples,path: " + spec, LG_ERR_INVALID);
                const int samples = std::atoi(f[4].substr(0, comma).c_str());
                string path = f[4].substr(comma + 1);
                for (size_t k = 5; k < f.size(); ++k) path += ":" + f[k];
                int w = 0, h = 0;
                const std::vector<float> rgb = readEXR_RGB(path, &w, &h);
                envMaps.emplace_back(new EnvMap(ctx, w, h, rgb.data()));
                double c[3];
                strToVec(f[2]).to(c);
                check_host(lg_light_env(std::stod(f[1]), c, std::stod(f[3]), samples, envMaps.back()->get(), &l),
                           "env light");
            } else if (lg_light_parse(spec.c_str(), &l) != LG_OK) {
                cout << spec << " is not a valid light: " << lg_host_last_error() << endl;
                continue;
            }
            opts.light.push_back(l);
        }

        std::pair<DenseBrickGrid, std::unique_ptr<DenseBrickGrid>> grids{DenseBrickGrid(ctx, nullptr), nullptr};
        if (!vdbFilename.empty()) {
            const VdbFloatGrid &d = vdb.density;
            const Vec3d origin(d.translation[0] + d.voxel_size * d.dense_min[0],
                               d.translation[1] + d.voxel_size * d.dense_min[1],
                               d.translation[2] + d.voxel_size * d.dense_min[2]);
            grids.first = DenseBrickGrid::fromDense(ctx, d.dims[0], d.dims[1], d.dims[2], d.dense.data(), d.voxel_size,
                                                    origin, spec.dtype);
            if (vdb.has_temperature)
                grids.second.reset(new DenseBrickGrid(DenseBrickGrid::fromDense(
                    ctx, d.dims[0], d.dims[1], d.dims[2], vdb.temperature.data(), d.voxel_size, origin, LG_F32)));
            if (opts.verbose) {
                cout << "Input Density Grid Voxel size: " << d.voxel_size << endl;
                cout << "Input Density Grid: " << d.dims[0] << "x" << d.dims[1] << "x" << d.dims[2] << " voxels from "
                     << d.leaves << " leaves, " << (d.empty ? "no active voxel" : "active box") << " [" << d.dense_min[0]
                     << "," << d.dense_min[1] << "," << d.dense_min[2] << "]" << endl;
                if (vdb.has_temperature)
                    cout << "Input Temperature Grid: " << (vdb.temperature_resampled ? "resampled" : "same transform")
                         << endl;
            }
            vdb = VdbVolume();  // the host copies are on the device now
        } else {
            cout << "Generating: " << scene << " " << sceneSize << "^3 (" << precision << " bricks)" << endl;
            grids = DenseBrickGrid::procedural(ctx, spec);
        }
        DenseBrickGrid &grid = grids.first;
        const DenseBrickGrid *tempGrid = grids.second.get();

        // COMPUTE LIGHTGRAPH ITERATIONS (main.cc:505-553)
        auto start1 = std::chrono::steady_clock::now();
        std::vector<std::vector<Vec3d>> pointCloudGather(opts.iter), diffusePointGather(opts.iter);
        LightGraph lg(grid, tempGrid, opts, shader, pointCloudGather, diffusePointGather);
        lg(0, opts.iter);
        std::vector<Vec3d> pointCloudCombined, diffusePointCloudCombined;
        for (size_t i = 0; i < opts.iter; ++i) {
            pointCloudCombined.insert(pointCloudCombined.end(), pointCloudGather[i].begin(), pointCloudGather[i].end());
            diffusePointCloudCombined.insert(diffusePointCloudCombined.end(), diffusePointGather[i].begin(),
                                             diffusePointGather[i].end());
        }
        ctx.sync();
        cout << "LightGraphs Calculation: " << secondsSince(start1) << " sec" << endl;
        if (opts.verbose) {
            cout << "Point Cloud size: " << pointCloudCombined.size() << endl;
            if (opts.sampler == "dart")
                cout << "Edge radius: 1.7 x each iteration's dart-throwing minimum distance" << endl;
            else
                cout << "Edge radius: " << lg.radius() << endl;
        }

        // RASTERIZE LIGHTGRAPH TO GRID (main.cc:560-589)
        auto start = std::chrono::steady_clock::now();
        ScatterGrid scatterGrid(grid, opts);
        scatterGrid.splat(pointCloudCombined, diffusePointCloudCombined, opts.ptsPerKernel * opts.iter);
        ctx.sync();
        cout << "Rasterize LightGraphs to Grid: " << secondsSince(start) << " sec" << endl;

        // RENDER (main.cc:72-128)
        auto start2 = std::chrono::steady_clock::now();
        Film rgbFilm(opts.width, opts.height), directFilm(opts.width, opts.height),
            scatterFilm(opts.width, opts.height), emissionFilm(opts.width, opts.height);
        VolumeRender renderer(grid, &scatterGrid, tempGrid, opts, shader);
        renderer.render(rgbFilm, directFilm, scatterFilm, emissionFilm);
        cout << "Render: " << secondsSince(start2) << " sec" << endl;
        cout << "Writting to Disk: " << imgFilename << "\n" << endl;
        saveEXR(imgFilename, rgbFilm, directFilm, scatterFilm, emissionFilm, opts);
    } catch (const std::exception &e) {
        std::cerr << "lightgraph: " << e.what() << endl;
        return EXIT_FAILURE;
    }
    return 0;
}
