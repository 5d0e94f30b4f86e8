// Test-only stand-in for the threshold-list types of calcQPI's qpi.h: one
// entry (sub-pixel i, j, orbitals o1, o2, factor) and the (2w+1)^4 nested
// lists indexed [xn1][yn1][xn2][yn2].
#pragma once
#include <cstddef>
#include <vector>

struct flaglistentry {
    size_t i, j, o1, o2;
    double factor;
};
typedef std::vector<std::vector<std::vector<std::vector<std::vector<flaglistentry> > > > > flaglist;
