// lattice.hpp — compile-time D2Q9 / D3Q19 tables for host and device code.
//
// The numbers come from lattice_tables.inc, generated by gen_lattice.py from
// the lattice definition (CLIP's WMRTvelSet order, include/solver/WMRT.cuh:189-206)
// and the moment polynomials of the weighted-MRT transform (WMRT.cuh:54-184).
#pragma once

namespace clipamd {

#include "lattice_tables.inc"

enum : int { SIDE_XM = 0, SIDE_XP = 1, SIDE_YM = 2, SIDE_YP = 3, SIDE_ZM = 4, SIDE_ZP = 5 };

template <int QQ>
struct Lattice;

template <>
struct Lattice<9> {
    static constexpr int Q = 9, D = 2, NSTRESS = 2, NIN = 3;
    __host__ __device__ static constexpr int ex(int q) { return D2Q9_Ex[q]; }
    __host__ __device__ static constexpr int ey(int q) { return D2Q9_Ey[q]; }
    __host__ __device__ static constexpr int ez(int q) { return D2Q9_Ez[q]; }
    __host__ __device__ static constexpr int e(int q, int a) { return a == 0 ? D2Q9_Ex[q] : (a == 1 ? D2Q9_Ey[q] : 0); }
    __host__ __device__ static constexpr double w(int q) { return D2Q9_W[q]; }
    __host__ __device__ static constexpr int opp(int q) { return D2Q9_OPP[q]; }
    __host__ __device__ static constexpr int M(int r, int q) { return D2Q9_M[r][q]; }
    __host__ __device__ static constexpr int minv_num(int q, int r) { return D2Q9_MINV_NUM[q][r]; }
    __host__ __device__ static constexpr int minv_den(int q) { return D2Q9_MINV_DEN[q]; }
    __host__ __device__ static constexpr int stress(int k) { return D2Q9_STRESS[k]; }
    // direction q mirrored along axis a
    __host__ __device__ static constexpr int mirror(int q, int a) { return D2Q9_MIRROR[a][q]; }
    // position of q in side `side`'s list of the reference's wall map /
    // slip map (WMRT.cuh:214-264; -1: q does not enter through that side)
    __host__ __device__ static constexpr int wall_pos(int side, int q) { return D2Q9_WALL_POS[side][q]; }
    __host__ __device__ static constexpr int slip_pos(int side, int q) { return D2Q9_SLIP_POS[side][q]; }
};

template <>
struct Lattice<19> {
    static constexpr int Q = 19, D = 3, NSTRESS = 5, NIN = 5;
    __host__ __device__ static constexpr int ex(int q) { return D3Q19_Ex[q]; }
    __host__ __device__ static constexpr int ey(int q) { return D3Q19_Ey[q]; }
    __host__ __device__ static constexpr int ez(int q) { return D3Q19_Ez[q]; }
    __host__ __device__ static constexpr int e(int q, int a) {
        return a == 0 ? D3Q19_Ex[q] : (a == 1 ? D3Q19_Ey[q] : D3Q19_Ez[q]);
    }
    __host__ __device__ static constexpr double w(int q) { return D3Q19_W[q]; }
    __host__ __device__ static constexpr int opp(int q) { return D3Q19_OPP[q]; }
    __host__ __device__ static constexpr int M(int r, int q) { return D3Q19_M[r][q]; }
    __host__ __device__ static constexpr int minv_num(int q, int r) { return D3Q19_MINV_NUM[q][r]; }
    __host__ __device__ static constexpr int minv_den(int q) { return D3Q19_MINV_DEN[q]; }
    __host__ __device__ static constexpr int stress(int k) { return D3Q19_STRESS[k]; }
    // direction q mirrored along axis a
    __host__ __device__ static constexpr int mirror(int q, int a) { return D3Q19_MIRROR[a][q]; }
    // position of q in side `side`'s list of the reference's wall map /
    // slip map (WMRT.cuh:214-264; -1: q does not enter through that side)
    __host__ __device__ static constexpr int wall_pos(int side, int q) { return D3Q19_WALL_POS[side][q]; }
    __host__ __device__ static constexpr int slip_pos(int side, int q) { return D3Q19_SLIP_POS[side][q]; }
};

using D2Q9 = Lattice<9>;
using D3Q19 = Lattice<19>;

}  // namespace clipamd
