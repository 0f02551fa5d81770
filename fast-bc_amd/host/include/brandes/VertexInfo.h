/*
 * VertexInfo (brandes/VertexInfo.h:17-571): per-vertex shortest-path lengths
 * (W) and counts (V) to the ascending borders of its cluster.  Same members
 * and semantics:
 *   - set/get with std::out_of_range at index >= borders();
 *   - normalize subtracts the minimum length, zeros of unreached borders
 *     included (VertexInfo.h:259-278);
 *   - squaredDistance sums, per border, the squared length difference and
 *     then the squared count difference (VertexInfo.h:293-307);
 *   - compare is lexicographic per border, count first, then length
 *     (VertexInfo.h:551-569);
 *   - element-wise arithmetic with another VertexInfo or a scalar.
 */
#ifndef FASTBC_AMD_VERTEXINFO_H
#define FASTBC_AMD_VERTEXINFO_H

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class VertexInfo {
public:
    explicit VertexInfo(int borderCount) : _n(borderCount), _len(borderCount), _cnt(borderCount) {}

    template <typename N, typename E>
    VertexInfo(const VertexInfo<N, E> &o) : _n(o._n), _len(o._n), _cnt(o._n)
    {
        for (int i = 0; i < _n; ++i) {
            _len[i] = (W)o._len[i];
            _cnt[i] = (V)o._cnt[i];
        }
    }

    VertexInfo(const VertexInfo &o) = default;

    template <typename N, typename E>
    VertexInfo &operator=(const VertexInfo<N, E> &o)
    {
        _n = o._n;
        _len.resize(_n);
        _cnt.resize(_n);
        for (int i = 0; i < _n; ++i) {
            _len[i] = (W)o._len[i];
            _cnt[i] = (V)o._cnt[i];
        }
        return *this;
    }

    VertexInfo &operator=(const VertexInfo &o) = default;

    void setBorderSPLength(int i, W length) { at(_len, i) = length; }
    W getBorderSPLength(int i) const { return at(_len, i); }
    void setBorderSPCount(int i, V count) { at(_cnt, i) = count; }
    V getBorderSPCount(int i) const { return at(_cnt, i); }

    W getMinBorderSPLength() const { return _n ? *std::min_element(_len.begin(), _len.end()) : W(0); }

    void normalize()
    {
        const W m = getMinBorderSPLength();
        for (int i = 0; i < _n; ++i) _len[i] -= m;
    }

    void reset()
    {
        _len.assign(_n, W(0));
        _cnt.assign(_n, V(0));
    }

    int borders() const { return _n; }

    template <typename N, typename E>
    W squaredDistance(const VertexInfo<N, E> &o) const
    {
        W sq = 0;
        for (int i = 0; i < _n; ++i) {
            const W dl = _len[i] - (W)o._len[i];
            sq += dl * dl;
            const W dc = (W)(_cnt[i] - (V)o._cnt[i]);
            sq += dc * dc;
        }
        return sq;
    }

#define FASTBC_VI_OP(op)                                                              \
    template <typename N, typename E>                                                 \
    VertexInfo &operator op##=(const VertexInfo<N, E> &o)                             \
    {                                                                                 \
        for (int i = 0; i < _n; ++i) {                                                \
            _len[i] op## = o._len[i];                                                 \
            _cnt[i] op## = o._cnt[i];                                                 \
        }                                                                             \
        return *this;                                                                 \
    }                                                                                 \
    template <typename N, typename E>                                                 \
    VertexInfo operator op(const VertexInfo<N, E> &o)                                 \
    {                                                                                 \
        VertexInfo r(*this);                                                          \
        return r op## = o;                                                            \
    }                                                                                 \
    template <typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
    VertexInfo &operator op##=(T num)                                                 \
    {                                                                                 \
        for (int i = 0; i < _n; ++i) {                                                \
            _len[i] op## = num;                                                       \
            _cnt[i] op## = num;                                                       \
        }                                                                             \
        return *this;                                                                 \
    }                                                                                 \
    template <typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type> \
    VertexInfo operator op(T num)                                                     \
    {                                                                                 \
        VertexInfo r(*this);                                                          \
        return r op## = num;                                                          \
    }
    FASTBC_VI_OP(+)
    FASTBC_VI_OP(-)
    FASTBC_VI_OP(*)
    FASTBC_VI_OP(/)
#undef FASTBC_VI_OP

    template <typename N, typename E>
    W compare(const VertexInfo<N, E> &o) const
    {
        for (int i = 0; i < _n; ++i) {
            if (W c = _cnt[i] - o._cnt[i]; c != 0) return c;
            if (W c = _len[i] - o._len[i]; c != 0) return c;
        }
        return 0;
    }
    template <typename N, typename E> bool operator==(const VertexInfo<N, E> &o) const { return compare(o) == 0; }
    template <typename N, typename E> bool operator!=(const VertexInfo<N, E> &o) const { return compare(o) != 0; }
    template <typename N, typename E> bool operator<(const VertexInfo<N, E> &o) const { return compare(o) < 0; }
    template <typename N, typename E> bool operator>(const VertexInfo<N, E> &o) const { return compare(o) > 0; }
    template <typename N, typename E> bool operator<=(const VertexInfo<N, E> &o) const { return compare(o) <= 0; }
    template <typename N, typename E> bool operator>=(const VertexInfo<N, E> &o) const { return compare(o) >= 0; }

private:
    template <typename T>
    T &at(std::vector<T> &v, int i)
    {
        if (i < 0 || i >= _n) throw std::out_of_range("Given store index is out of range.");
        return v[i];
    }
    template <typename T>
    const T &at(const std::vector<T> &v, int i) const
    {
        if (i < 0 || i >= _n) throw std::out_of_range("Given store index is out of range.");
        return v[i];
    }

    int _n;
    std::vector<W> _len;
    std::vector<V> _cnt;

    template <typename N, typename E>
    friend class VertexInfo;
};

} // namespace brandes
} // namespace fastbc

#endif
