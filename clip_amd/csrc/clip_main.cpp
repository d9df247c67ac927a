// clip_main.cpp — the `clip` executable, the reference driver (src/main.cu:39-91)
// written against include/clip_amd.hpp: read ./config.txt, build the host
// classes around a DataArray, fresh start or `-resume`, then the step loop
// with VTS output, reporting and checkpoints.  Differences: fields
// are copied to the host only on output steps (the reference copies them
// every step, VTSwriter.cu:72-79); optional `-precision f64|f32|mixed_f16|
// mixed_bf16`, `-config <path>`, `-steps <n>` and `-novts` flags.
#include "clip_amd.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

int main(int argc, char** argv) {
    bool resume = false;
    std::string config = "config.txt";
    int precision = CLIP_PREC_F64;
    long long steps_override = -1;
    bool vts = true;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a == "-resume") resume = true;
        else if (a == "-config" && i + 1 < argc) config = argv[++i];
        else if (a == "-steps" && i + 1 < argc) steps_override = std::atoll(argv[++i]);
        else if (a == "-novts") vts = false;
        else if (a == "-precision" && i + 1 < argc) {
            const std::string p(argv[++i]);
            precision = (p == "f32") ? CLIP_PREC_F32 : (p == "mixed_f16") ? CLIP_PREC_MIXED_F16
                      : (p == "mixed_bf16") ? CLIP_PREC_MIXED_BF16 : CLIP_PREC_F64;
        }
    }
    try {
        clip::InputData inputData(config);
        clip::Domain simDomain(inputData);
        clip::Geometry geometry(inputData);
        clip::Boundary boundaryConditions(inputData, simDomain);
        clip::DataArray dataArray(inputData, simDomain, boundaryConditions);
        dataArray.createVectors();

        clip::NSAllen nsSolver(inputData, simDomain, dataArray, boundaryConditions, geometry, precision);
        clip::TimeInfo timeInfo(inputData);
        clip::VTSwriter vtsWriter(dataArray, inputData, simDomain, timeInfo, "results", "results");
        clip::Reporter reporter(dataArray, inputData, simDomain, timeInfo);
        clip::CheckPointer checkpointer(dataArray, inputData, simDomain, timeInfo, boundaryConditions);
        clip::NSAllen& ns = nsSolver;

        if (resume) {
            checkpointer.load();
            nsSolver.macroscopic();  // already part of the load: skipped
        } else {
            nsSolver.initialCondition();
            dataArray.updateDevice();
            nsSolver.deviceInitializer();
        }
        if (vts) vtsWriter.writeToFile();
        const long long final_step = steps_override >= 0 ? timeInfo.getCurrentStep() + steps_override
                                                         : (long long)timeInfo.getFinalStep();
        const auto t0 = std::chrono::steady_clock::now();
        long long done = 0;
        while (timeInfo.getCurrentStep() < final_step) {
            ns.solve();
            timeInfo.increment();
            ++done;
            if (vts) vtsWriter.writeToFile();
            reporter.print();
            checkpointer.save();
        }
        clip::check(clip_synchronize(ns.handle()));
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        clip_info inf;
        clip::check(clip_get_info(ns.handle(), &inf));
        if (done > 0)
            std::printf("steps %lld  time %.3f s  MLUPS %.1f\n", done, sec, (double)inf.interior * done / sec / 1e6);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
