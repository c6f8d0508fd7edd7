// Khuzdul application API on MI355X: integer types (include/types.h of the
// reference, same names and widths).
#ifndef TYPES_H
#define TYPES_H

#include <stdint.h>

typedef uint32_t VertexId;
typedef uint64_t EdgeId;
typedef uint32_t LabelId;
typedef uint8_t PVertexId;
typedef uint8_t PEdgeId;

#endif
