// elas_cli.cpp -- command-line front end of the ELAS path, mirroring the
// reference's demo program (src/Main.cu:36-134):
//
//   elas_cli demo [image_dir]        the four demo pairs (teddy, cones, aloe,
//                                    raindeer) from image_dir (default img)
//   elas_cli left.pgm right.pgm      one stereo pair
//   elas_cli -h                      usage
//
// Options (before or after the positional arguments):
//   -o PREFIX        output prefix of a single pair (default: the left
//                    image's path without ".pgm"); demo outputs go next to
//                    PREFIX/<name> when -o is given, else to the current dir
//   --disp-max N     disparity range 0..N (default: the reference's 255)
//
// The reference shows D1 in an OpenCV window (Main.cu:14-34).  This tool
// writes <prefix>_D1.pgm / _D2.pgm with the same 8-bit scaling and
// <prefix>_D1.pfm / _D2.pfm with the float maps.  Like the reference it
// prints the device properties first (Main.cu:36-66) and reports pairs of
// different sizes instead of processing them.  The exit status is 0 when
// every pair was processed and written, 1 otherwise.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "elas/Elas.h"

namespace {

void print_device_props() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) != hipSuccess) continue;
        printf("Device Number: %d\n", i);
        printf("  Device name: %s (%s)\n", p.name, p.gcnArchName);
        printf("  Memory Clock Rate (KHz): %d\n", p.memoryClockRate);
        printf("  Memory Bus Width (bits): %d\n", p.memoryBusWidth);
        printf("  Peak Memory Bandwidth (GB/s): %f\n", 2.0 * p.memoryClockRate * (p.memoryBusWidth / 8) / 1.0e6);
        printf("  Compute Units: %d\n", p.multiProcessorCount);
        printf("  Max Threads Per Block: %d\n", p.maxThreadsPerBlock);
        printf("  Shared Mem Per Block: %zu bytes\n\n", p.sharedMemPerBlock);
    }
}

void usage() {
    printf("\nELAS demo program usage:\n");
    printf("./elas_cli demo [dir] ........... process all test images (image dir, default img)\n");
    printf("./elas_cli left.pgm right.pgm ... process a single stereo pair\n");
    printf("./elas_cli -h ................... shows this help\n");
    printf("options: -o PREFIX (output prefix), --disp-max N (default 255)\n\n");
    printf("Note: All images must be pgm greylevel images. Outputs are <prefix>_D1/_D2.pgm\n");
    printf("      (disparities scaled such that the largest = 255) and <prefix>_D1/_D2.pfm.\n\n");
}

bool write_maps(const std::string& prefix, const float* D, int W, int H, const char* tag) {
    std::vector<uint8_t> u8((size_t)W * H);
    elas_io_disparity_to_u8(D, u8.data(), (int64_t)W * H);
    const std::string pgm = prefix + "_" + tag + ".pgm", pfm = prefix + "_" + tag + ".pfm";
    const bool ok = elas_io_write_pgm(pgm.c_str(), u8.data(), W, H) == ELAS_OK &&
                    elas_io_write_pfm(pfm.c_str(), D, W, H) == ELAS_OK;
    if (!ok) printf("ERROR: could not write %s / %s\n", pgm.c_str(), pfm.c_str());
    return ok;
}

// Main.cu:68-98
bool process(Elas& elas, const char* file_1, const char* file_2, const std::string& prefix) {
    printf("Processing: %s, %s\n", file_1, file_2);
    int32_t w1 = 0, h1 = 0, w2 = 0, h2 = 0;
    if (elas_io_read_pgm(file_1, nullptr, 0, &w1, &h1) != ELAS_OK ||
        elas_io_read_pgm(file_2, nullptr, 0, &w2, &h2) != ELAS_OK) {
        printf("ERROR: Could not read file %s or %s\n", file_1, file_2);
        return false;
    }
    if (w1 != w2 || h1 != h2) {
        printf("ERROR: Images must be of same size, but\n");
        printf("       I1: %d x %d, I2: %d x %d\n", w1, h1, w2, h2);
        return false;
    }
    const size_t n = (size_t)w1 * h1;
    std::vector<uint8_t> I1(n), I2(n);
    if (elas_io_read_pgm(file_1, I1.data(), (int64_t)n, &w1, &h1) != ELAS_OK ||
        elas_io_read_pgm(file_2, I2.data(), (int64_t)n, &w2, &h2) != ELAS_OK) {
        printf("ERROR: Could not read file %s or %s\n", file_1, file_2);
        return false;
    }
    const int32_t dims[2] = {w1, h1};
    std::vector<float> D1(n), D2(n);
    elas.process(I1.data(), I2.data(), D1.data(), D2.data(), dims);
    if (elas.status() != ELAS_OK) {
        printf("ERROR: %s (status %d)\n", elas.last_error(), elas.status());
        if (elas.status() != ELAS_ERR_TOO_FEW_POINTS) return false;
    }
    bool ok = write_maps(prefix, D1.data(), w1, h1, "D1");
    ok = write_maps(prefix, D2.data(), w1, h1, "D2") && ok;
    if (ok) printf("  wrote %s_D1/_D2 .pgm/.pfm\n", prefix.c_str());
    return ok && elas.status() == ELAS_OK;
}

std::string strip_pgm(const std::string& s) {
    return (s.size() > 4 && s.compare(s.size() - 4, 4, ".pgm") == 0) ? s.substr(0, s.size() - 4) : s;
}

}  // namespace

int main(int argc, char* argv[]) {
    std::vector<std::string> pos;
    std::string out;
    Elas::parameters param;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "-o" && i + 1 < argc) {
            out = argv[++i];
        } else if (a == "--disp-max" && i + 1 < argc) {
            param.disp_max = atoi(argv[++i]);
        } else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            pos.push_back(a);
        }
    }
    if (!((pos.size() == 1 || pos.size() == 2) && pos[0] == "demo") && pos.size() != 2) {
        usage();
        return pos.empty() ? 0 : 1;
    }
    print_device_props();
    Elas elas(param, 0);
    if (!elas.context()) {
        printf("ERROR: no usable gfx950 device (status %d)\n", elas.status());
        return 1;
    }
    bool ok = true;
    if (pos[0] == "demo") {
        const std::string dir = pos.size() == 2 ? pos[1] : std::string("img");
        for (const char* name : {"teddy", "cones", "aloe", "raindeer"}) {
            const std::string l = dir + "/" + name + "_left.pgm", r = dir + "/" + name + "_right.pgm";
            const std::string prefix = (out.empty() ? std::string(".") : out) + "/" + name;
            ok = process(elas, l.c_str(), r.c_str(), prefix) && ok;
        }
    } else {
        ok = process(elas, pos[0].c_str(), pos[1].c_str(), out.empty() ? strip_pgm(pos[0]) : out);
    }
    printf("... done!\n");
    return ok ? 0 : 1;
}
