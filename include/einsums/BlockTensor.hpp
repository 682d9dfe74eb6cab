// Block-sparse device tensors and their einsums as grouped persistent launches.
//
// Reference types mirrored: BlockDeviceTensor (libs/Einsums/Tensor/include/Einsums/Tensor/
// BlockTensor.hpp:46,119-131: hypercubic blocks, one extent per block) and
// TiledDeviceTensor (TiledTensor.hpp:56,93-133: per-axis tile extents, a map from tile
// coordinates to tiles, absent tiles are zero). einsum on them follows BlockAlgebra.hpp:23-52
// (block i from blocks i) and TileAlgebra.hpp:18-93 (C scaled once, then every tile triple of
// the tile grid accumulates with prefactor 1); all block contractions of one call are handed
// to the backend as one grouped plan (ein_grouped_plan_*), i.e. persistent launches instead of
// one library call per block.
#pragma once

#include <algorithm>
#include <map>
#include <vector>

#include "TensorAlgebra.hpp"

namespace einsums {

template <typename T, size_t Rank>
class BlockDeviceTensor {
  public:
    using ValueType              = T;
    static constexpr size_t Rank_ = Rank;

    BlockDeviceTensor(std::string name, std::vector<int64_t> block_dims) : name_(std::move(name)), dims_(std::move(block_dims)) {
        for (int64_t n : dims_) {
            std::array<int64_t, Rank> d;
            d.fill(n);
            blocks_.emplace_back(name_ + " block", d);
        }
    }
    size_t num_blocks() const { return blocks_.size(); }
    int64_t block_dim(size_t i) const { return dims_[i]; }
    DeviceTensor<T, Rank> &block(size_t i) { return blocks_[i]; }
    DeviceTensor<T, Rank> const &block(size_t i) const { return blocks_[i]; }
    DeviceTensor<T, Rank> &operator[](size_t i) { return blocks_[i]; }
    void zero() {
        for (auto &b : blocks_) b.zero();
    }
    std::string const &name() const { return name_; }

  private:
    std::string name_;
    std::vector<int64_t> dims_;
    std::vector<DeviceTensor<T, Rank>> blocks_;
};

template <typename T, size_t Rank>
class TiledDeviceTensor {
  public:
    using ValueType = T;
    using Coord     = std::array<int, Rank>;

    TiledDeviceTensor(std::string name, std::array<std::vector<int64_t>, Rank> tile_dims)
        : name_(std::move(name)), tile_dims_(std::move(tile_dims)) {}

    size_t grid_size(size_t d) const { return tile_dims_[d].size(); }
    std::array<int64_t, Rank> tile_shape(Coord const &c) const {
        std::array<int64_t, Rank> s;
        for (size_t d = 0; d < Rank; ++d) s[d] = tile_dims_[d][size_t(c[d])];
        return s;
    }
    DeviceTensor<T, Rank> &tile(Coord const &c) {
        auto it = tiles_.find(c);
        if (it == tiles_.end()) it = tiles_.emplace(c, DeviceTensor<T, Rank>(name_ + " tile", tile_shape(c))).first;
        return it->second;
    }
    bool has_tile(Coord const &c) const {
        auto it = tiles_.find(c);
        return it != tiles_.end() && it->second.size() > 0;
    }
    DeviceTensor<T, Rank> const &tile(Coord const &c) const { return tiles_.at(c); }
    std::map<Coord, DeviceTensor<T, Rank>> &tiles() { return tiles_; }

  private:
    std::string name_;
    std::array<std::vector<int64_t>, Rank> tile_dims_;
    std::map<Coord, DeviceTensor<T, Rank>> tiles_;
};

// RAII handle of a backend grouped plan (reusable, capturable into a hipGraph).
class GroupedPlan {
  public:
    GroupedPlan() = default;
    GroupedPlan(int elem_bytes, std::vector<ein_contraction_item> const &items) {
        check(ein_grouped_plan_create(elem_bytes, items.data(), int32_t(items.size()), &plan_), "grouped plan");
    }
    GroupedPlan(GroupedPlan &&o) noexcept : plan_(o.plan_) { o.plan_ = nullptr; }
    GroupedPlan &operator=(GroupedPlan &&o) noexcept {
        std::swap(plan_, o.plan_);
        return *this;
    }
    GroupedPlan(GroupedPlan const &)            = delete;
    GroupedPlan &operator=(GroupedPlan const &) = delete;
    ~GroupedPlan() {
        if (plan_) (void)ein_grouped_plan_destroy(plan_);
    }
    void execute() const { check(ein_grouped_plan_execute(plan_, gpu::stream()), "grouped execute"); }
    void autotune() const { check(ein_grouped_plan_autotune(plan_, gpu::stream()), "grouped autotune"); }
    int waves() const {
        int32_t rows = 0, w = 0;
        int64_t t = 0;
        check(ein_grouped_plan_info(plan_, &rows, &w, &t), "grouped info");
        return int(w);
    }

  private:
    void *plan_ = nullptr;
};

namespace tensor_algebra {
namespace detail {

template <typename TT, size_t R>
ein_tensor_desc block_desc(DeviceTensor<TT, R> const &t, int32_t const *labels) {
    return t.desc(labels);
}

} // namespace detail

// Builds the per-block grouped plan of a BlockDeviceTensor einsum (BlockAlgebra.hpp:23-52).
template <typename T, size_t CR, size_t AR, size_t BR, typename... CIndices, typename... AIndices, typename... BIndices>
GroupedPlan make_block_plan(T C_prefactor, std::tuple<CIndices...> const &, BlockDeviceTensor<T, CR> *C, T AB_prefactor,
                            std::tuple<AIndices...> const &, BlockDeviceTensor<T, AR> const &A,
                            std::tuple<BIndices...> const &, BlockDeviceTensor<T, BR> const &B) {
    static_assert(sizeof...(CIndices) == CR && sizeof...(AIndices) == AR && sizeof...(BIndices) == BR, "rank mismatch");
    if (C->num_blocks() != A.num_blocks() || C->num_blocks() != B.num_blocks())
        throw tensor_compat_error("block einsum: block counts differ");
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    std::vector<ein_contraction_item> items;
    for (size_t i = 0; i < C->num_blocks(); ++i) {
        if (C->block_dim(i) != A.block_dim(i) || C->block_dim(i) != B.block_dim(i))
            throw dimension_error("block einsum: block extents differ");
        if (C->block_dim(i) == 0) continue;
        ein_contraction_item it{};
        it.C            = C->block(i).desc(c_ids.data());
        it.A            = A.block(i).desc(a_ids.data());
        it.B            = B.block(i).desc(b_ids.data());
        it.c            = C->block(i).data();
        it.a            = A.block(i).data();
        it.b            = B.block(i).data();
        it.c_prefactor  = double(C_prefactor);
        it.ab_prefactor = double(AB_prefactor);
        items.push_back(it);
    }
    return GroupedPlan(int(sizeof(T)), items);
}

template <typename T, size_t CR, size_t AR, size_t BR, typename U, typename... CIndices, typename... AIndices,
          typename... BIndices>
void einsum(U const C_prefactor, std::tuple<CIndices...> const &Ci, BlockDeviceTensor<T, CR> *C, U const AB_prefactor,
            std::tuple<AIndices...> const &Ai, BlockDeviceTensor<T, AR> const &A, std::tuple<BIndices...> const &Bi,
            BlockDeviceTensor<T, BR> const &B) {
    make_block_plan(T(C_prefactor), Ci, C, T(AB_prefactor), Ai, A, Bi, B).execute();
}

// TiledDeviceTensor einsum (TileAlgebra.hpp:18-93): C scaled once by C_prefactor, then every
// tile triple over the tile grid of the distinct labels accumulates with prefactor 1;
// triples with an absent or empty tile are skipped. Triples writing the same C tile run in
// order (separate waves of the plan).
template <typename T, size_t CR, size_t AR, size_t BR, typename... CIndices, typename... AIndices, typename... BIndices>
GroupedPlan make_tiled_plan(T AB_prefactor, std::tuple<CIndices...> const &, TiledDeviceTensor<T, CR> *C,
                            std::tuple<AIndices...> const &, TiledDeviceTensor<T, AR> &A, std::tuple<BIndices...> const &,
                            TiledDeviceTensor<T, BR> &B) {
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    std::vector<int32_t> labels;
    std::vector<size_t> grid;
    auto add = [&](int32_t id, size_t n) {
        auto it = std::find(labels.begin(), labels.end(), id);
        if (it == labels.end()) {
            labels.push_back(id);
            grid.push_back(n);
        } else if (grid[size_t(it - labels.begin())] != n) {
            throw dimension_error("tiled einsum: tile grids disagree");
        }
    };
    for (size_t d = 0; d < CR; ++d) add(c_ids[d], C->grid_size(d));
    for (size_t d = 0; d < AR; ++d) add(a_ids[d], A.grid_size(d));
    for (size_t d = 0; d < BR; ++d) add(b_ids[d], B.grid_size(d));
    auto pos_of = [&](int32_t id) { return size_t(std::find(labels.begin(), labels.end(), id) - labels.begin()); };
    std::vector<ein_contraction_item> items;
    std::vector<size_t> combo(labels.size(), 0);
    size_t total = 1;
    for (auto n : grid) total *= n;
    for (size_t lin = 0; lin < total; ++lin) {
        size_t rem = lin;
        for (size_t q = labels.size(); q-- > 0;) {
            combo[q] = rem % grid[q];
            rem /= grid[q];
        }
        typename TiledDeviceTensor<T, CR>::Coord cc;
        typename TiledDeviceTensor<T, AR>::Coord ac;
        typename TiledDeviceTensor<T, BR>::Coord bc;
        for (size_t d = 0; d < CR; ++d) cc[d] = int(combo[pos_of(c_ids[d])]);
        for (size_t d = 0; d < AR; ++d) ac[d] = int(combo[pos_of(a_ids[d])]);
        for (size_t d = 0; d < BR; ++d) bc[d] = int(combo[pos_of(b_ids[d])]);
        if (!C->has_tile(cc) || !A.has_tile(ac) || !B.has_tile(bc)) continue;
        ein_contraction_item it{};
        it.C            = C->tile(cc).desc(c_ids.data());
        it.A            = A.tile(ac).desc(a_ids.data());
        it.B            = B.tile(bc).desc(b_ids.data());
        it.c            = C->tile(cc).data();
        it.a            = A.tile(ac).data();
        it.b            = B.tile(bc).data();
        it.c_prefactor  = 1.0;
        it.ab_prefactor = double(AB_prefactor);
        items.push_back(it);
    }
    return GroupedPlan(int(sizeof(T)), items);
}

template <typename T, size_t CR, size_t AR, size_t BR, typename U, typename... CIndices, typename... AIndices,
          typename... BIndices>
void einsum(U const C_prefactor, std::tuple<CIndices...> const &Ci, TiledDeviceTensor<T, CR> *C, U const AB_prefactor,
            std::tuple<AIndices...> const &Ai, TiledDeviceTensor<T, AR> &A, std::tuple<BIndices...> const &Bi,
            TiledDeviceTensor<T, BR> &B) {
    for (auto &kv : C->tiles()) kv.second.scale(T(C_prefactor));
    make_tiled_plan(T(AB_prefactor), Ci, C, Ai, A, Bi, B).execute();
}

} // namespace tensor_algebra
} // namespace einsums
