// The design handle behind ktplace_gen.h's opaque ktg_design, shared by the
// generator / flattening (netgen.cc) and the Bookshelf reader (bookshelf.cc).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "ktplace_gen.h"

struct ktg_design {
    // raw
    std::vector<uint8_t> isFixed;
    std::vector<double> posX, posY, width, height;
    std::vector<uint64_t> netStart;
    std::vector<uint32_t> pinVertex;
    std::vector<double> pinOffX, pinOffY, netWeight;
    std::vector<double> rowY, rowH, rowX0, rowX1, rowPitch;
    double die[4] = {0, 0, 0, 0};
    uint32_t numMovable = 0;
    // flat
    bool flat = false;
    std::vector<uint64_t> fNetStart;
    std::vector<uint32_t> fPinCell;
    std::vector<double> fOffX, fOffY, fWeight;
    std::vector<uint32_t> varOfVertex, movVertex;
    std::vector<double> fixedX, fixedY;
    std::vector<uint64_t> pinStart;
    std::vector<uint32_t> cellPinNet, cellPinIdx;
    std::vector<double> mWidth, mHeight, mArea;
    std::vector<double> obstacles;
    // vertex names (read designs only; empty for generated ones)
    std::vector<std::string> names;
    std::string nameBlob;  // names joined by '\n', built on request
};


namespace ktgi {

// Records msg as the calling thread's last error and returns -1.
int fail(const std::string &msg);
const char *lastErrorText();

// Flattens raw into d (collectCells / collectNets, simpl_dm.cc:16-85); d also
// keeps a copy of the raw arrays.
int collect(const ktg_raw &raw, ktg_design &d);

}  // namespace ktgi
