// tsu_datadat.cpp -- the reference's runtime configuration file (data.dat).
//
// Format (data.dat:1-45, read by Problem_loader.cxx:264-335): a first
// comment line, then for each value one comment line starting with '#'
// followed by a line holding the value.  Values, in order: xmin, xmax, ymin,
// ymax, total time, save time, npx, npy, CFL, epsilon_h, R, DS, friction,
// porosity, critical theta, Dsg, output prefix, bathymetry file, eta/CC
// file, fixed-bed file, sea-level correction, tides file.
//
// Derived parameters follow Problem_loader.cxx:337-342 (quadratic friction
// law: the friction coefficient is divided by 8) and the grid spacing of the
// analytic-input path (Problem_loader.cxx:631-632).
#define TSU_BUILDING_LIBRARY 1
#include "../../include/tsuner.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "tsu_host.h"

namespace {

void copy_str(char* dst, int len, const std::string& s)
{
    if (!dst || len <= 0) return;
    std::strncpy(dst, s.c_str(), (size_t)len - 1);
    dst[len - 1] = '\0';
}

}  // namespace

extern "C" {

TSU_API int tsu_parse_data_dat(const char* path, tsu_sim_params* params, char* prefix, int prefix_len,
                               char* files, int files_len, double* sea_level_correction)
{
    if (!path || !params) return tsu::set_error(TSU_ERR_ARG, "null argument");
    std::ifstream f(path);
    if (!f) return tsu::set_error(TSU_ERR_HOST, std::string("cannot open ") + path);
    std::vector<std::string> tok;
    std::string line;
    bool first = true;
    while (std::getline(f, line)) {
        if (first) {   // leading comment line
            first = false;
            continue;
        }
        size_t k = line.find_first_not_of(" \t\r");
        if (k == std::string::npos || line[k] == '#') continue;
        std::istringstream ss(line);
        std::string t;
        if (ss >> t) tok.push_back(t);
    }
    if (tok.size() < 22) return tsu::set_error(TSU_ERR_ARG, "data.dat holds " + std::to_string(tok.size()) + " values, 22 expected");
    tsu_sim_params p;
    std::memset(&p, 0, sizeof(p));
    try {
        p.xmin = std::stod(tok[0]);
        p.xmax = std::stod(tok[1]);
        p.ymin = std::stod(tok[2]);
        p.ymax = std::stod(tok[3]);
        p.total_time = std::stod(tok[4]);
        p.save_time = std::stod(tok[5]);
        p.npx = std::stoi(tok[6]);
        p.npy = std::stoi(tok[7]);
        p.CFL = std::stod(tok[8]);
        p.epsilon_h = std::stod(tok[9]);
        p.R = std::stod(tok[10]);
        p.DS = std::stod(tok[11]);
        p.FR = std::stod(tok[12]);
        p.porosity = std::stod(tok[13]);
        p.OC = std::stod(tok[14]);
        p.Dsg = std::stod(tok[15]);
        if (sea_level_correction) *sea_level_correction = std::stod(tok[20]);
    } catch (...) {
        return tsu::set_error(TSU_ERR_ARG, "malformed number in data.dat");
    }
    if (p.npx < 1 || p.npy < 1) return tsu::set_error(TSU_ERR_ARG, "npx and npy must be positive");
    const double G = 9.81;
    // Problem_loader.cxx:337-342
    p.FR /= 8.0;
    p.HQB = 5.0 * p.DS;
    p.Q = p.DS * std::sqrt(G * (p.R - 1.0) * p.DS) / (1.0 - p.porosity);
    p.Opar = p.FR / ((p.R - 1.0) * G * p.DS);
    p.nvol = p.npx * p.npy;
    p.vmax = 10.0;
    // Problem_loader.cxx:631-632
    p.dx = (p.xmax - p.xmin) / (p.npx);
    p.dy = (p.ymax - p.ymin) / (p.npy);
    *params = p;
    copy_str(prefix, prefix_len, tok[16]);
    // bathymetry, eta/CC, fixed-bed and tides files, separated by '\n'
    copy_str(files, files_len, tok[17] + "\n" + tok[18] + "\n" + tok[19] + "\n" + tok[21]);
    return TSU_OK;
}

}  // extern "C"
