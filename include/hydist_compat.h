/*
 * hydist_compat.h - the Hydist_C host surface (engine.h, loader.h,
 * main.cpp's init sequence) implemented on the gfx950 library.
 *
 * A program written against the reference's engine.h / loader.h keeps its
 * calls: load_inputs, attribute_arrays_memory_alloc, supporting_arrays_alloc,
 * Load_coeffs, Hydraulic_Calculation, save_result, save_FS.  The kernel
 * launches of main.cpp:257-299 (Onetime_init, Find_Calculation_limits_*,
 * Htuongdoi, preprocess_data) become one call, hydist_init.
 *
 * Differences a caller can observe:
 *  - Argument_Pointers' field pointers are device buffers of the library's
 *    context in the reference's dense layout (width M+3); the segment
 *    tables daui/cuoii/dauj/cuoij use a per-line capacity of (len+3)/2+2
 *    instead of segment_limit (no 20-segment cap).
 *  - Errors are printed and abort, as synch_and_check does
 *    (engine.cpp:13-19); the C ABI (hydist.h) returns codes instead.
 *  - load_inputs loads each series into the parameter named for it.  A
 *    caller that keeps main.cpp:138's argument order (bc_down passed in
 *    the bc_right slot) therefore gets the reference's swapped sides.
 *  - Boundary types are read per side, boundary_type[0..3] = up, down,
 *    left, right (the reference reads past a 4-entry array, Q7).
 */
#ifndef HYDIST_COMPAT_H
#define HYDIST_COMPAT_H

#include <string>
#include <utility>
#include <vector>

#include "hydist.h"

#define DOUBLE double
#define segment_limit 20

/* Options (engine.h:16-37). */
struct Options {
    DOUBLE Tmax, t_start, sediment_start, bed_change_start;
    int interval;
    bool kenhhepd, kenhhepng;
    int M, N;
    bool ketdinh, channel, debug, plot;
    int total_time;
    Options() = default;
    Options(DOUBLE Tmx, DOUBLE ts, int itv, DOUBLE sds, DOUBLE bcs, bool cohesive, bool kenhng, bool kenhd, bool db,
            bool plt)
        : Tmax(Tmx), t_start(ts), sediment_start(sds), bed_change_start(bcs), interval(itv), kenhhepd(kenhd),
          kenhhepng(kenhng), M(0), N(0), ketdinh(cohesive), channel(kenhd ^ kenhng), debug(db), plot(plt),
          total_time(0) {}
};

/* Constant_Coeffs (engine.h:39-52): the same 49 doubles. */
typedef hd_coeffs Constant_Coeffs;

/* Argument_Pointers (engine.h:55-69). */
struct Argument_Pointers {
    int M, N;
    DOUBLE hmax_u, hmax_d, hmax_l, hmax_r;
    int *bienQ;
    int *daui, *dauj, *cuoii, *cuoij, *moci, *mocj, *khouot, *boundary_type;
    DOUBLE *h, *v, *u, *z, *t_u, *t_v, *t_z, *Htdu, *Htdv, *H_moi, *htaiz, *htaiz_bd;
    DOUBLE *ubt, *ubp, *vbt, *vbd;
    DOUBLE *hsnham, *VISCOIDX, *Kx1, *Ky1, *Tsyw, *Tsxw;
    DOUBLE *bc_up, *bc_down, *bc_left, *bc_right;
    DOUBLE *hi;
    DOUBLE *FS, *tFS, *CC_u, *CC_d, *CC_l, *CC_r;
    DOUBLE *VTH, *Kx, *Ky, *Fw;
    DOUBLE *Qbx, *Qby;
    DOUBLE *dH;
};

/* Array_Pointers (engine.h:71-78): the sweep scratch of the context. */
struct Array_Pointers {
    DOUBLE *a1, *b1, *c1, *d1, *a2, *c2, *d2;
    DOUBLE *f1, *f2, *f3, *f5;
    DOUBLE *AA, *BB, *CC, *DD;
    DOUBLE *x;
    DOUBLE *Ap, *Bp, *ep;
    int *SN;
};

/* Host_arrays (engine.h:80-88). */
struct Host_arrays {
    int M, N, total_time;
    std::vector<DOUBLE> h, hsnham, VISCOIDX, Fw, FS, dH;
    std::vector<DOUBLE> bc_up, bc_down, bc_left, bc_right;
    std::vector<DOUBLE> CC_u, CC_d, CC_l, CC_r;
    std::vector<DOUBLE> u, v, z;
    std::vector<int> khouot, bienQ, boundary_type;
};

/* loader.cpp:4-87 */
void Load_coeffs(Constant_Coeffs &ce);

/* loader.cpp:91-112: whitespace text grid; returns (rows, cols). */
template <typename T>
std::pair<int, int> load_file(const char *filename, std::vector<T> &arr);

/* loader.cpp:120-168, same parameter order (bc_up, bc_right, bc_left,
 * bc_down).  Each series is loaded into the parameter it is named for. */
std::pair<int, int> load_inputs(std::string dir, std::vector<DOUBLE> &h, std::vector<DOUBLE> &hsnham,
                                std::vector<DOUBLE> &VISCOIDX, std::vector<DOUBLE> &Fw,
                                std::vector<DOUBLE> &bc_up, std::vector<DOUBLE> &bc_right,
                                std::vector<DOUBLE> &bc_left, std::vector<DOUBLE> &bc_down,
                                std::vector<DOUBLE> &CC_u, std::vector<DOUBLE> &CC_d, std::vector<DOUBLE> &CC_l,
                                std::vector<DOUBLE> &CC_r, std::vector<int> &bienQ, std::vector<int> &boundary_type,
                                int *total_time);

/* loader.cpp:224-330: allocates and uploads one scenario; *device_arg_ptr
 * receives a device copy of the returned struct (the key of the context). */
Argument_Pointers attribute_arrays_memory_alloc(int device, Host_arrays &ap, Argument_Pointers **device_arg_ptr);

/* loader.cpp:333-375 */
Array_Pointers supporting_arrays_alloc(int M, int N, Array_Pointers **device_arr_ptr);

/* main.cpp:257-299: Onetime_init, segmentation, Htuongdoi, preprocess_data.
 * coeffs may be a host or a device pointer. */
void hydist_init(Argument_Pointers *d_arg_ptr, Constant_Coeffs *coeffs);

/* engine.cpp:22-36 */
void update_boundary_at_t(int M, int N, float t, bool channel, int total_time, Argument_Pointers *d_arg_ptr,
                          Constant_Coeffs *coeffs);

/* engine.cpp:13-19 */
void synch_and_check();

/* engine.cpp:41-296 */
void Hydraulic_Calculation(DOUBLE dT, DOUBLE NANGDAY, Argument_Pointers *d_arg_ptr, Array_Pointers *d_arr_ptr,
                           Constant_Coeffs *coeffs, Argument_Pointers h_arg_ptr, Options ops);

/* loader.cpp:397-434: text output under Outputs/ of the working directory. */
void save_result(Argument_Pointers h_arg_pointer, int t);
void save_FS(Argument_Pointers h_arg_pointer, int t, DOUBLE ros);

/* The hd_ctx behind a compat Argument_Pointers (device or host copy). */
hd_ctx *hydist_compat_context(const Argument_Pointers *arg);

#endif
