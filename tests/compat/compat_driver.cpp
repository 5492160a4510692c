// compat_driver.cpp - a host program written against the reference's
// engine.h / loader.h surface (include/hydist_compat.h), in the order of the
// reference's main(): load_inputs, the two allocators, Load_coeffs, the init
// sequence and Hydraulic_Calculation.  Flags follow main.cpp:58-110:
//   -f DIR  input directory (trailing '/')   -h/-m/-s  run length
//   -i N    save interval (s)                -b T      sediment start (s)
//   -B T    bed-change start (s)
//   -d FILE after the run, write Z, U, V (full (N+3) x (M+3) grids of
//           doubles, in that order) to FILE for comparisons at full
//           precision
// Outputs land in ./Outputs/{U,V,Z,FS}/ as in loader.cpp:397-434.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "hydist_compat.h"

int main(int argc, char **argv) {
    std::string dir, dump;
    int hour = 1, minute = 0, sec = 0, save_interval = 60;
    DOUBLE sediment_start = 20 * 3600, bed_change_start = 500 * 3600;
    int c;
    while ((c = getopt(argc, argv, "f:h:m:s:b:B:i:d:")) != -1) {
        switch (c) {
            case 'f': dir = optarg; break;
            case 'h': hour = std::atoi(optarg); break;
            case 'm': minute = std::atoi(optarg); break;
            case 's': sec = std::atoi(optarg); break;
            case 'b': sediment_start = std::atoi(optarg); break;
            case 'B': bed_change_start = std::atoi(optarg); break;
            case 'i': save_interval = std::atoi(optarg); break;
            case 'd': dump = optarg; break;
            default: return 2;
        }
    }
    const DOUBLE Tmax = hour * 3600 + minute * 60 + sec;
    Options ops(Tmax, 0.0, save_interval, sediment_start, bed_change_start, true, false, false, false, false);

    std::vector<DOUBLE> h, hsnham, VISCOIDX, Fw, bc_up, bc_down, bc_left, bc_right, CC_u, CC_d, CC_l, CC_r;
    std::vector<int> bienQ, boundary_type;
    int total_time = 0;
    // the reference's call order (bc_down in the bc_right slot, main.cpp:138)
    std::pair<int, int> size = load_inputs(dir, h, hsnham, VISCOIDX, Fw, bc_up, bc_down, bc_left, bc_right, CC_u,
                                           CC_d, CC_l, CC_r, bienQ, boundary_type, &total_time);
    const int N = size.first - 3, M = size.second - 3;
    ops.M = M;
    ops.N = N;
    ops.total_time = total_time;

    Host_arrays host;
    host.M = M;
    host.N = N;
    host.total_time = total_time;
    host.h = h;
    host.hsnham = hsnham;
    host.VISCOIDX = VISCOIDX;
    host.Fw = Fw;
    host.bc_up = bc_up;
    host.bc_down = bc_down;
    host.bc_left = bc_left;
    host.bc_right = bc_right;
    host.CC_u = CC_u;
    host.CC_d = CC_d;
    host.CC_l = CC_l;
    host.CC_r = CC_r;
    host.bienQ = bienQ;
    host.boundary_type = boundary_type;

    Argument_Pointers *d_arg = nullptr;
    Array_Pointers *d_arr = nullptr;
    Argument_Pointers h_arg = attribute_arrays_memory_alloc(0, host, &d_arg);
    supporting_arrays_alloc(M, N, &d_arr);
    Constant_Coeffs coeffs;
    Load_coeffs(coeffs);
    hydist_init(d_arg, &coeffs);
    synch_and_check();
    std::cout << "mesh " << N << " x " << M << ", total_time " << total_time << std::endl;
    Hydraulic_Calculation(coeffs.dT, coeffs.NANGDAY, d_arg, d_arr, &coeffs, h_arg, ops);
    synch_and_check();
    std::cout << "done t = " << Tmax << std::endl;
    if (!dump.empty()) {
        hd_ctx *ctx = hydist_compat_context(&h_arg);
        const size_t n = (size_t)(M + 3) * (N + 3);
        std::vector<double> buf(n);
        FILE *f = std::fopen(dump.c_str(), "wb");
        if (!ctx || !f) return 3;
        for (int field : {HD_F_Z, HD_F_U, HD_F_V}) {
            if (hd_copy_out(ctx, field, buf.data(), n) != HD_OK) return 3;
            if (std::fwrite(buf.data(), sizeof(double), n, f) != n) return 3;
        }
        if (std::fclose(f) != 0) return 3;
    }
    return 0;
}
