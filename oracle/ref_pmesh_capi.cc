// C entry point onto the reference's own SerialMesh and ParallelMesh
// classes -- TEST INFRASTRUCTURE ONLY.  Compiled together with the
// reference sources SerialMesh.cc and ParallelMesh.cc (which build on their
// own with PARALLEL_MESH_READ_SERIAL_ONLY) into oracle/_ref/libref_pmesh.so
// by `make -C oracle ref`; the tests compare the package's .pmesh writer
// and reader with it.
#include <cstdint>
#include <string>
#include <vector>

#include "ParallelMesh.hh"
#include "SerialMesh.hh"

// Partition the serial mesh in smesh_path cell by cell as partition[] says
// and write it as a parallel mesh to pmesh_path.  Returns the cell count.
extern "C" int64_t ref_smesh_to_pmesh(const char *smesh_path,
                                      const char *pmesh_path,
                                      const uint64_t *partition,
                                      int64_t n_cells, int n_parts)
{
    SerialMesh serial;
    serial.read(std::string(smesh_path));
    if (static_cast<int64_t>(serial.c_numCells) != n_cells)
        return -1;
    std::vector<uint64_t> part(partition, partition + n_cells);
    ParallelMesh parallel;
    parallel.createFromSerialMesh(serial, part, n_parts);
    parallel.write(std::string(pmesh_path));
    return n_cells;
}
