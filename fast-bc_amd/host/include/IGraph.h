/*
 * Graph interface of the reference (IGraph.h:15-62), kept with the same
 * names and signatures so that code written against libfastbc compiles here.
 */
#ifndef FASTBC_AMD_IGRAPH_H
#define FASTBC_AMD_IGRAPH_H

#include <map>
#include <vector>

namespace fastbc {

template <typename V, typename W>
class IGraph {
public:
    virtual ~IGraph() = default;
    /* weight of src->dest, 0 when absent */
    virtual W edge(V src, V dest) const = 0;
    /* outgoing arcs of src, by ascending target */
    virtual const std::map<V, W> &forwardStar(V src) const = 0;
    /* incoming arcs of dest, by ascending source */
    virtual const std::map<V, W> &backwardStar(V dest) const = 0;
    /* vertex ids 0..n-1 */
    virtual const std::vector<V> &vertices() const = 0;
    /* arc count */
    virtual V edges() const = 0;
};

} // namespace fastbc

#endif
