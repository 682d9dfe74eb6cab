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
#include <cstdint>
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
    // Calls f(coordinate) for every present tile of non-zero size.
    template <typename F>
    void for_each_tile(F &&f) const {
        for (auto const &kv : tiles_)
            if (kv.second.size() > 0) f(kv.first);
    }

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

template <typename X>
inline constexpr bool is_block_v = false;
template <typename T, size_t R>
inline constexpr bool is_block_v<BlockDeviceTensor<T, R>> = true;
template <typename X>
inline constexpr bool is_tiled_v = false;
template <typename T, size_t R>
inline constexpr bool is_tiled_v<TiledDeviceTensor<T, R>> = true;

template <typename TT, size_t R>
ein_tensor_desc block_desc(DeviceTensor<TT, R> const &t, int32_t const *labels) {
    return t.desc(labels);
}

// The sub-block [off, off+n) on every axis of a dense device tensor: the reference's
// std::apply(T, view_index) with every Range equal to one block's range.
template <typename T, size_t R>
DeviceTensorView<T, R> diag_view(DeviceTensorBase<T, R> const &t, int64_t off, int64_t n) {
    std::array<int64_t, R> d;
    d.fill(n);
    int64_t o = 0;
    for (size_t i = 0; i < R; ++i) o += off * t.stride(i);
    return DeviceTensorView<T, R>(t.name() + " block", t.data() + o, d, t.strides());
}

inline ein_contraction_item make_item(ein_tensor_desc const &C, void *c, ein_tensor_desc const &A, void const *a,
                                      ein_tensor_desc const &B, void const *b, double cp, double abp) {
    ein_contraction_item it{};
    it.C            = C;
    it.A            = A;
    it.B            = B;
    it.c            = c;
    it.a            = a;
    it.b            = b;
    it.c_prefactor  = cp;
    it.ab_prefactor = abp;
    return it;
}

// Scalar C of a block or tiled einsum (BlockAlgebra.hpp:105-180, TileAlgebra.hpp:95-158):
// every item writes its own partial sum, then C = C_prefactor*C + (partials added in item
// order) by the deterministic dot. The result is read back (a synchronisation).
template <typename T, typename CType>
void scalar_from_items(T C_prefactor, CType *C, std::vector<ein_contraction_item> items) {
    size_t const nt = items.size();
    DeviceTensor<T, 1> partial("block partial sums", int64_t(std::max<size_t>(nt, 1)));
    DeviceTensor<T, 1> ones("ones", std::array<int64_t, 1>{int64_t(std::max<size_t>(nt, 1))},
                            std::vector<T>(std::max<size_t>(nt, 1), T(1)));
    for (size_t k = 0; k < nt; ++k) {
        items[k].C           = ein_tensor_desc{};
        items[k].c           = partial.data() + k;
        items[k].c_prefactor = 0.0;
    }
    if (nt) GroupedPlan(int(sizeof(T)), items).execute();
    DeviceTensor<T, 0> c("scalar C");
    if (C_prefactor != T(0)) c = T(*C);
    int st;
    if constexpr (std::is_same_v<T, double>)
        st = ein_dot_f64(int64_t(nt), 1.0, partial.data(), 1, ones.data(), 1, double(C_prefactor), c.data(), gpu::stream());
    else
        st = ein_dot_f32(int64_t(nt), 1.0f, partial.data(), 1, ones.data(), 1, float(C_prefactor), c.data(), gpu::stream());
    check(st, "block scalar sum");
    *C = CType(T(c));
}

// Items of an einsum over BlockDeviceTensors where A and/or B is a BlockDeviceTensor and the
// other may be a dense device tensor cut into the diagonal sub-block of every block's range;
// C is a BlockDeviceTensor (block i), a dense device tensor (its diagonal sub-block i; the
// rest untouched) or a scalar (placeholder items). Blocks of extent 0 are skipped
// (BlockAlgebra.hpp:23-428).
template <typename T, typename CType, typename AType, typename BType, size_t NC, size_t NA, size_t NB>
std::vector<ein_contraction_item> block_items(T C_prefactor, std::array<int32_t, NC> const &c_ids, CType *C,
                                              T AB_prefactor, std::array<int32_t, NA> const &a_ids, AType const &A,
                                              std::array<int32_t, NB> const &b_ids, BType const &B) {
    static_assert(is_block_v<AType> || is_block_v<BType>, "A or B must be a BlockDeviceTensor");
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>,
                  "block and tiled einsums run on double and float (one grouped launch of real items)");
    auto const &blk = [&]() -> auto const & {
        if constexpr (is_block_v<AType>) return A;
        else return B;
    }();
    size_t const nb = blk.num_blocks();
    auto check_blocks = [&](auto const &X) {
        if constexpr (is_block_v<std::remove_cvref_t<decltype(X)>>) {
            if (X.num_blocks() != nb) throw tensor_compat_error("block einsum: block counts differ");
            for (size_t i = 0; i < nb; ++i)
                if (X.block_dim(i) != blk.block_dim(i)) throw dimension_error("block einsum: block extents differ");
        }
    };
    check_blocks(A);
    check_blocks(B);
    if constexpr (is_block_v<CType>) check_blocks(*C);
    std::vector<ein_contraction_item> items;
    int64_t off = 0;
    for (size_t i = 0; i < nb; ++i) {
        int64_t const n = blk.block_dim(i);
        if (n == 0) continue;
        auto part = [&](auto const &X, auto const &ids) {
            using X_t = std::remove_cvref_t<decltype(X)>;
            if constexpr (is_block_v<X_t>) return X.block(i).view();
            else return diag_view(X, off, n);
        };
        auto av = part(A, a_ids);
        auto bv = part(B, b_ids);
        ein_tensor_desc cd{};
        void *cptr = nullptr;
        if constexpr (is_block_v<CType>) {
            cd   = C->block(i).desc(c_ids.data());
            cptr = C->block(i).data();
        } else if constexpr (!std::is_arithmetic_v<CType>) {
            auto cv = diag_view(*C, off, n);
            cd      = cv.desc(c_ids.data());
            cptr    = cv.data();
        }
        items.push_back(make_item(cd, cptr, av.desc(a_ids.data()), av.data(), bv.desc(b_ids.data()), bv.data(),
                                  double(C_prefactor), double(AB_prefactor)));
        off += n;
    }
    return items;
}

// A BlockDeviceTensor seen as a tiled tensor whose tiles are its diagonal blocks
// (BlockTileAlgebra.hpp:18-173).
template <typename T, size_t R>
struct BlockTiles {
    BlockDeviceTensor<T, R> const *b;
    using Coord = std::array<int, R>;
    size_t grid_size(size_t) const { return b->num_blocks(); }
    bool has_tile(Coord const &c) const {
        for (size_t d = 1; d < R; ++d)
            if (c[d] != c[0]) return false;
        return b->block_dim(size_t(c[0])) > 0;
    }
    DeviceTensor<T, R> const &tile(Coord const &c) const { return b->block(size_t(c[0])); }
    template <typename F>
    void for_each_tile(F &&f) const {
        for (size_t i = 0; i < b->num_blocks(); ++i)
            if (b->block_dim(i) > 0) {
                Coord c;
                c.fill(int(i));
                f(c);
            }
    }
};

template <typename T, size_t R>
auto as_tiles(TiledDeviceTensor<T, R> &t) -> TiledDeviceTensor<T, R> & {
    return t;
}
template <typename T, size_t R>
auto as_tiles(TiledDeviceTensor<T, R> const &t) -> TiledDeviceTensor<T, R> & {
    return const_cast<TiledDeviceTensor<T, R> &>(t);
}
template <typename T, size_t R>
BlockTiles<T, R> as_tiles(BlockDeviceTensor<T, R> const &b) {
    return BlockTiles<T, R>{&b};
}

} // namespace detail

// Builds the per-block grouped plan of a BlockDeviceTensor einsum (BlockAlgebra.hpp:23-52).
template <typename T, size_t CR, size_t AR, size_t BR, typename... CIndices, typename... AIndices, typename... BIndices>
GroupedPlan make_block_plan(T C_prefactor, std::tuple<CIndices...> const &, BlockDeviceTensor<T, CR> *C, T AB_prefactor,
                            std::tuple<AIndices...> const &, BlockDeviceTensor<T, AR> const &A,
                            std::tuple<BIndices...> const &, BlockDeviceTensor<T, BR> const &B) {
    static_assert(sizeof...(CIndices) == CR && sizeof...(AIndices) == AR && sizeof...(BIndices) == BR, "rank mismatch");
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    return GroupedPlan(int(sizeof(T)), detail::block_items(C_prefactor, c_ids, C, AB_prefactor, a_ids, A, b_ids, B));
}

// Einsums with BlockDeviceTensor operands, every combination of the reference's special
// dispatch (BlockAlgebra.hpp:23-428): A and B blocks, or one of them a dense device tensor
// (cut by every block's range); C a BlockDeviceTensor, a dense device tensor (its diagonal
// blocks) or a scalar. One grouped plan per call.
template <typename U, typename CType, typename AType, typename BType, typename... CIndices, typename... AIndices,
          typename... BIndices>
    requires((detail::is_block_v<AType> || detail::is_block_v<BType>) &&
             (detail::is_block_v<AType> || detail::DeviceTensorConcept<AType>) &&
             (detail::is_block_v<BType> || detail::DeviceTensorConcept<BType>) &&
             (detail::is_block_v<CType> || detail::DeviceTensorConcept<CType> || std::is_arithmetic_v<CType>))
void einsum(U const C_prefactor, std::tuple<CIndices...> const &, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &, AType const &A, std::tuple<BIndices...> const &, BType const &B) {
    using T = std::conditional_t<detail::is_block_v<AType>, typename AType::ValueType, typename BType::ValueType>;
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    auto items = detail::block_items(T(C_prefactor), c_ids, C, T(AB_prefactor), a_ids, A, b_ids, B);
    if constexpr (std::is_arithmetic_v<CType>) {
        detail::scalar_from_items(T(C_prefactor), C, std::move(items));
    } else if (!items.empty()) {
        GroupedPlan(int(sizeof(T)), items).execute();
    }
}

template <typename X>
struct is_block_v_helper : std::bool_constant<detail::is_block_v<X>> {};

// Tile-triple plan of an einsum over TiledDeviceTensors, or over a mix of TiledDeviceTensors
// and BlockDeviceTensors (TileAlgebra.hpp:18-93, BlockTileAlgebra.hpp:18-173): every tile
// triple over the tile grid of the distinct labels accumulates with prefactor 1; triples with
// an absent or empty A or B tile are skipped. A TiledDeviceTensor C gets a zero tile where a
// triple lands on an absent one (the reference's C->tile(index)); a BlockDeviceTensor C has
// only its diagonal positions. Triples writing the same C tile run in order (separate waves
// of the plan). With C == nullptr the items carry no C (a scalar C, placed by the caller).
template <typename T, typename CT, typename AT, typename BT, size_t NC, size_t NA, size_t NB>
std::vector<ein_contraction_item> tile_items(T AB_prefactor, std::array<int32_t, NC> const &c_ids, CT *C,
                                             std::array<int32_t, NA> const &a_ids, AT const &A,
                                             std::array<int32_t, NB> const &b_ids, BT const &B) {
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
    if (C != nullptr)
        for (size_t d = 0; d < NC; ++d) add(c_ids[d], C->grid_size(d));
    for (size_t d = 0; d < NA; ++d) add(a_ids[d], A.grid_size(d));
    for (size_t d = 0; d < NB; ++d) add(b_ids[d], B.grid_size(d));
    auto pos_of = [&](int32_t id) { return size_t(std::find(labels.begin(), labels.end(), id) - labels.begin()); };
    // Join of the present A and B tiles on their shared labels (no walk over the whole tile
    // grid: with irrep-blocked tensors most positions have no tile), then the labels found only
    // in C; the triples are sorted into the order in which the grid walk of
    // TileAlgebra.hpp:59-92 meets them, so every C tile accumulates in the reference's order.
    std::vector<std::array<int, NA>> a_tiles;
    std::vector<std::array<int, NB>> b_tiles;
    A.for_each_tile([&](std::array<int, NA> const &c) { a_tiles.push_back(c); });
    B.for_each_tile([&](std::array<int, NB> const &c) { b_tiles.push_back(c); });
    std::vector<size_t> c_only;
    for (size_t d = 0; d < NC; ++d)
        if (C != nullptr && std::find(a_ids.begin(), a_ids.end(), c_ids[d]) == a_ids.end() &&
            std::find(b_ids.begin(), b_ids.end(), c_ids[d]) == b_ids.end() &&
            std::find(c_only.begin(), c_only.end(), pos_of(c_ids[d])) == c_only.end())
            c_only.push_back(pos_of(c_ids[d]));
    std::vector<std::vector<size_t>> combos;
    for (auto const &ac : a_tiles) {
        std::vector<size_t> combo(labels.size(), SIZE_MAX);
        bool ok = true;
        auto set = [&](size_t q, size_t v) {
            if (combo[q] != SIZE_MAX && combo[q] != v) ok = false;
            combo[q] = v;
        };
        for (size_t d = 0; d < NA; ++d) set(pos_of(a_ids[d]), size_t(ac[d]));
        if (!ok) continue;
        for (auto const &bc : b_tiles) {
            std::vector<size_t> cb = combo;
            bool okb = true;
            for (size_t d = 0; d < NB; ++d) {
                size_t const q = pos_of(b_ids[d]);
                if (cb[q] != SIZE_MAX && cb[q] != size_t(bc[d])) okb = false;
                cb[q] = size_t(bc[d]);
            }
            if (!okb) continue;
            size_t extra = 1;
            for (size_t q : c_only) extra *= grid[q];
            for (size_t e = 0; e < extra; ++e) {
                size_t rem = e;
                for (size_t x = c_only.size(); x-- > 0;) {
                    cb[c_only[x]] = rem % grid[c_only[x]];
                    rem /= grid[c_only[x]];
                }
                combos.push_back(cb);
            }
        }
    }
    std::sort(combos.begin(), combos.end());
    std::vector<ein_contraction_item> items;
    for (auto const &combo : combos) {
        std::array<int, NA> ac;
        std::array<int, NB> bc;
        for (size_t d = 0; d < NA; ++d) ac[d] = int(combo[pos_of(a_ids[d])]);
        for (size_t d = 0; d < NB; ++d) bc[d] = int(combo[pos_of(b_ids[d])]);
        auto const &ta = A.tile(ac);
        auto const &tb = B.tile(bc);
        ein_tensor_desc cd{};
        void *cptr = nullptr;
        if (C != nullptr) {
            std::array<int, NC> cc;
            for (size_t d = 0; d < NC; ++d) cc[d] = int(combo[pos_of(c_ids[d])]);
            if constexpr (is_block_v_helper<CT>::value) {
                bool diag = true;
                for (size_t d = 1; d < NC; ++d) diag = diag && cc[d] == cc[0];
                if (!diag || C->block_dim(size_t(cc[0])) == 0) continue;
                cd   = C->block(size_t(cc[0])).desc(c_ids.data());
                cptr = C->block(size_t(cc[0])).data();
            } else {
                auto shape = C->tile_shape(cc);
                bool empty = false;
                for (auto e : shape) empty = empty || e == 0;
                if (empty) continue;
                bool const fresh = !C->has_tile(cc);
                auto &tc         = C->tile(cc); // created when absent
                if (fresh) tc.zero();
                cd = tc.desc(c_ids.data());
                cptr = tc.data();
            }
        }
        items.push_back(detail::make_item(cd, cptr, ta.desc(a_ids.data()), ta.data(), tb.desc(b_ids.data()), tb.data(), 1.0,
                                          double(AB_prefactor)));
    }
    return items;
}

template <typename T, size_t CR, size_t AR, size_t BR, typename... CIndices, typename... AIndices, typename... BIndices>
GroupedPlan make_tiled_plan(T AB_prefactor, std::tuple<CIndices...> const &, TiledDeviceTensor<T, CR> *C,
                            std::tuple<AIndices...> const &, TiledDeviceTensor<T, AR> &A, std::tuple<BIndices...> const &,
                            TiledDeviceTensor<T, BR> &B) {
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    return GroupedPlan(int(sizeof(T)), tile_items(AB_prefactor, c_ids, C, a_ids, A, b_ids, B));
}

// Tiled and mixed block/tiled einsums (TileAlgebra.hpp:18-158, BlockTileAlgebra.hpp:18-173):
// C (tiled or block) scaled once by C_prefactor (zeroed when it is 0), then every tile triple
// accumulates; a scalar C receives C_prefactor*C + the sum of the triples' contractions.
template <typename U, typename CType, typename AType, typename BType, typename... CIndices, typename... AIndices,
          typename... BIndices>
    requires((detail::is_tiled_v<AType> || detail::is_tiled_v<BType> || detail::is_tiled_v<CType>) &&
             (detail::is_tiled_v<AType> || detail::is_block_v<AType>) &&
             (detail::is_tiled_v<BType> || detail::is_block_v<BType>) &&
             (detail::is_tiled_v<CType> || detail::is_block_v<CType> || std::is_arithmetic_v<CType>))
void einsum(U const C_prefactor, std::tuple<CIndices...> const &, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &, AType const &A, std::tuple<BIndices...> const &, BType const &B) {
    using T = typename AType::ValueType;
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    auto At = detail::as_tiles(A);
    auto Bt = detail::as_tiles(B);
    if constexpr (std::is_arithmetic_v<CType>) {
        using NoC = TiledDeviceTensor<T, 1>;
        auto items = tile_items<T, NoC>(T(AB_prefactor), std::array<int32_t, 1>{0}, static_cast<NoC *>(nullptr), a_ids, At,
                                        b_ids, Bt);
        detail::scalar_from_items(T(C_prefactor), C, std::move(items));
    } else {
        if constexpr (detail::is_tiled_v<CType>) {
            for (auto &kv : C->tiles()) kv.second.scale(T(C_prefactor));
        } else {
            for (size_t i = 0; i < C->num_blocks(); ++i) C->block(i).scale(T(C_prefactor));
        }
        auto items = tile_items(T(AB_prefactor), c_ids, C, a_ids, At, b_ids, Bt);
        if (!items.empty()) GroupedPlan(int(sizeof(T)), items).execute();
    }
}

} // namespace tensor_algebra
} // namespace einsums
