#pragma once
#include <stdint.h>

#include <vector>

namespace mccfr {

// Preflop bucket of every ordered hole pair, [h0*52 + h1], 0 on the diagonal.
void build_preflop_lut(uint8_t* lut_2704);

// Arithmetic evaluator tables: non-flush class by rank-multiset index
// (49,205 entries), flush class by 13-bit rank mask (8,192), and the
// lexicographic-index offsets [13][8][5].
void build_arith_tables(std::vector<uint16_t>& nonflush, std::vector<uint16_t>& flush,
                        std::vector<uint16_t>& offsets);

}  // namespace mccfr
