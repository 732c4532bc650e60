// tsu_cli.cpp -- the `tsuner data.dat [device]` program (main_TSUNER.cxx:22-107).
//
// Reads data.dat, builds the initial state, runs TSUNER() (libtsuner) and
// writes <prefix>.nc, with the reference's messages and exit status.
// Initial state, in order of precedence:
//   --state FILE   raw little-endian binary: 5*npx*npy doubles (h, qx, qy,
//                  H, hc planes) followed by npx*npy int32 fixed-bed flags;
//   input files    when the bathymetry file named in data.dat exists: the
//                  reference's Read_files = 1 build (Problem_loader.cxx:355-624)
//                  through tsu_load_input_files, with the boundary series
//                  CC/tCC passed on to TSUNER();
//   otherwise      the reference's analytic example case (cini.cxx:22-61),
//                  sampled at cell centres as Problem_loader.cxx:650-685 does.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/tsuner.h"

namespace {

// cini.cxx:22-31: depth below datum, positive down.
double example_bathy(double x, double y)
{
    double H = -2.5 * std::exp(-5.0 * ((x + 0) * (x + 0) + (y - 0) * (y - 0)));
    H += -x * (x > 0.0);
    return H;
}

// cini.cxx:34-42
double example_h(double x, double y)
{
    double h1 = example_bathy(x, y) + 2.0 + 6 * std::exp(-5.0 * ((x + 2) * (x + 2) + (y - 0) * (y - 0)));
    return h1 * (h1 > 0.0);
}

bool file_exists(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: tsuner data.dat [device] [--state FILE]" << std::endl;
        return EXIT_FAILURE;
    }
    std::string input = argv[1];
    int device = 0;
    std::string state_file;
    for (int a = 2; a < argc; a++) {
        if (std::strcmp(argv[a], "--state") == 0 && a + 1 < argc)
            state_file = argv[++a];
        else
            device = std::atoi(argv[a]);
    }
    if (!file_exists(input)) {
        std::cerr << "Error: Folder not found. '" << input << "'" << std::endl;
        return EXIT_FAILURE;
    }
    switch (checkCUDAsupport(device)) {
    case 1:
        std::cerr << "Error: There is no GPU" << std::endl;
        return EXIT_FAILURE;
    case 2:
        std::cerr << "Error: There is no supported GPU (gfx950)" << std::endl;
        return EXIT_FAILURE;
    case 3:
        std::cerr << "Error: bad device index" << std::endl;
        return EXIT_FAILURE;
    }
    initializeGPU();

    tsu_sim_params p;
    char prefix[1024], files[4096];
    double slr = 0.0;
    if (tsu_parse_data_dat(input.c_str(), &p, prefix, sizeof(prefix), files, sizeof(files), &slr) != TSU_OK) {
        std::cerr << "Error: " << tsu_last_error() << std::endl;
        return EXIT_FAILURE;
    }
    const size_t nv = (size_t)p.npx * p.npy;
    std::vector<double> h(nv), H(nv);
    std::vector<tsu_double2> q(nv);
    std::vector<int> fix(nv, 1);
    std::vector<double> CC, tCC;
    std::vector<std::string> names;   // bathymetry, eta/CC, fixed bed, tides
    {
        std::string all(files), one;
        for (char ch : all + "\n") {
            if (ch == '\n') {
                names.push_back(one);
                one.clear();
            } else {
                one += ch;
            }
        }
    }
    if (state_file.empty() && names.size() == 4 && file_exists(names[0])) {
        if (names[3] != "0") {
            std::cerr << "Error: tides input ('" << names[3] << "') is not supported" << std::endl;
            return EXIT_FAILURE;
        }
        std::printf("Reading bathymetry, initial and boundary conditions\n");
        std::fflush(stdout);
        unsigned long T = 0;
        if (tsu_input_time_levels(names[1].c_str(), &T) != TSU_OK) {
            std::cerr << "Error: " << tsu_last_error() << std::endl;
            return EXIT_FAILURE;
        }
        CC.assign((size_t)(2 * p.npx + 2 * p.npy) * 4 * T, 0.0);
        tCC.assign(T, 0.0);
        if (tsu_load_input_files(names[0].c_str(), names[1].c_str(), names[2].c_str(), slr, &p, h.data(), q.data(),
                                 H.data(), fix.data(), CC.data(), tCC.data(), T) != TSU_OK) {
            std::cerr << "Error: " << tsu_last_error() << std::endl;
            return EXIT_FAILURE;
        }
    } else if (!state_file.empty()) {
        std::ifstream f(state_file, std::ios::binary);
        std::vector<double> V(5 * nv);
        if (!f.read((char*)V.data(), (std::streamsize)(V.size() * sizeof(double))) ||
            !f.read((char*)fix.data(), (std::streamsize)(nv * sizeof(int)))) {
            std::cerr << "Error: cannot read " << nv << " cells from '" << state_file << "'" << std::endl;
            return EXIT_FAILURE;
        }
        for (size_t i = 0; i < nv; i++) {
            h[i] = V[i];
            q[i].x = V[nv + i];
            q[i].y = V[2 * nv + i];
            H[i] = V[3 * nv + i];
        }
    } else {
        // Problem_loader.cxx:650-685 with the cini.cxx callbacks
        for (int j = 0; j < p.npy; j++)
            for (int i = 0; i < p.npx; i++) {
                const size_t pos = (size_t)j * p.npx + i;
                double x = p.xmin + 0.5 * p.dx + i * p.dx;
                double y = p.ymin + 0.5 * p.dy + j * p.dy;
                h[pos] = example_h(x, y);
                H[pos] = example_bathy(x, y);
                q[pos].x = 0.0 * h[pos];
                q[pos].y = 0.0 * h[pos];
            }
    }
    std::printf("%i\n", p.nvol);
    double seconds = 0.0;
    int err = TSUNER(h.data(), q.data(), H.data(), fix.data(), p, prefix, nullptr, CC.empty() ? nullptr : CC.data(),
                     tCC.empty() ? nullptr : tCC.data(), (unsigned long)tCC.size(), nullptr, nullptr, &seconds, device,
                     nullptr, nullptr, 0);
    if (err > 0) {
        if (err == 1)
            std::cerr << "Error: Not enough GPU memory" << std::endl;
        else
            std::cerr << "Error: Not enough CPU memory" << std::endl;
        return EXIT_FAILURE;
    }
    std::cout << std::scientific << std::endl << "Total simulation time: " << seconds << " sec." << std::endl;
    std::cout << "Normal End." << std::endl;
    return EXIT_SUCCESS;
}
