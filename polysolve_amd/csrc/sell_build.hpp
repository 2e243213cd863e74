// Bank-padded SELL-64 matrix built on the GPU from a column-compressed matrix in device memory.
#pragma once

#include "common.hpp"

#include <cstdint>
#include <vector>

namespace psmas
{
    /// The matrix in the solver's working space, as the host path builds it
    /// (host_setup.cpp build_padded_bsr, then the SELL-64 layout of Solver::upload_matrix).
    struct SellDevice
    {
        long long nnzb = 0;   ///< real blocks (the pattern, without SELL padding)
        long long nslots = 0; ///< SELL slots (64 rows each)
        std::vector<int> slice_ptr;         ///< host copy, nslices + 1
        std::vector<char> nonisolated;      ///< slot has a block off its diagonal
        DeviceBuffer<int> d_slice_ptr;      ///< nslices + 1
        DeviceBuffer<int> d_cols;           ///< nslots * 64
        DeviceBuffer<double> d_vals;        ///< nslots * 64 * bd * bd
        DeviceBuffer<int> d_row_len;        ///< npad: blocks per row
        DeviceBuffer<int> d_map;            ///< per input non-zero: its index in d_vals (when asked for)
        bool map_ok = false;                ///< d_map valid: no duplicate entries and int32 indices
    };

    /// The checks of validate_csc (host_setup.hpp) on CSC arrays in device memory (d_colptr of
    /// n+1 and d_rowidx of nnz entries, as the caller states); throws InvalidArgumentError with
    /// the same messages.  build_sell_device relies on them for its index arithmetic.
    void validate_csc_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx, hipStream_t s);

    /// Build from n x n CSC arrays in device memory (d_colptr n+1, d_rowidx and d_vals nnz) with
    /// block size bd, the block -> slot map d_slot_of_block (div_up(n, bd) entries) and npad slots.
    /// Same pattern and values as the host build: unique (block row, block column) slot pairs,
    /// columns ascending in every row, the trailing partial block's diagonal block with 1.0 on its
    /// padded diagonal, duplicate entries summed (in an unspecified order), every 64-row slice at
    /// least one slot long, padding slots holding the row's own column and zero values.  With
    /// want_map, also the value index of every non-zero (valid only without duplicate entries).
    void build_sell_device(int n, long long nnz, int bd, const int *d_colptr, const int *d_rowidx,
                           const double *d_vals, const int *d_slot_of_block, int npad, bool want_map,
                           hipStream_t s, SellDevice &out);

    /// map[keep[k]] for every k (-1 where keep[k] < 0): the value index of each non-zero of a
    /// matrix whose non-zero k became non-zero keep[k] of the built one (or was dropped).
    DeviceBuffer<int> compose_value_map(const std::vector<int> &keep, const DeviceBuffer<int> &map, hipStream_t s);
    /// The same with keep (n entries) in device memory.
    DeviceBuffer<int> compose_value_map(const DeviceBuffer<int> &keep, size_t n, const DeviceBuffer<int> &map,
                                        hipStream_t s);

    /// The block-diagonal matrix of the systems parts[p] (each built with npads[p] slots), one
    /// after another as the concatenated host build lays them out: system p's rows and columns
    /// moved by the slots of the systems before it, its slices after theirs.  Frees parts.
    void concat_sell(std::vector<SellDevice> &parts, const std::vector<int> &npads, hipStream_t s, SellDevice &out);
} // namespace psmas
