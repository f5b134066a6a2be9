"""Drop-in for pyaxisymflow/utils/dump_vtk.py:5-62 without the vtk package:
the same vtk_init / vtk_write calls write the same VTK XML ImageData file
(.vti): a grid_size_z x grid_size_r x 1 image of unit spacing whose point
data is one float64 array named "softy" with one named component per field,
fields flattened in C order (the reference's field.reshape(-1) copied into
the component).  The array is stored as base64 binary with a UInt32 byte
count header, the vtkXMLImageDataWriter default without compression."""
import base64
import struct

import numpy as np


class _ImageData:
    def __init__(self, grid_size_z, grid_size_r):
        self.dims = (int(grid_size_z), int(grid_size_r), 1)


class _Writer:
    def __init__(self, image):
        self.image = image
        self.filename = None

    def SetFileName(self, filename):
        self.filename = filename


def vtk_init(grid_size_z, grid_size_r):
    """dump_vtk.py:5-26: (image data, temporary array, writer)."""
    image = _ImageData(grid_size_z, grid_size_r)
    return image, None, _Writer(image)


def vtk_write(filename, vtk_image_data, temp_vtk_array, writer, field_names_list, field_list,
              grid_size_z, grid_size_r):
    """dump_vtk.py:29-62: every field in field_list becomes one component of
    the point array "softy", in order.  Fields may be numpy arrays or device
    tensors (copied to the host)."""
    n = int(grid_size_z) * int(grid_size_r)
    cols = []
    for f in field_list:
        if hasattr(f, "detach"):
            f = f.detach().cpu().numpy()
        a = np.asarray(f, dtype=np.float64).reshape(-1)
        if a.size != n:
            raise ValueError(f"field of {a.size} values for a {grid_size_z} x {grid_size_r} image")
        cols.append(a)
    data = np.stack(cols, axis=1) if cols else np.zeros((n, 0))
    raw = np.ascontiguousarray(data, dtype="<f8").tobytes()
    payload = base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")
    nz, nr = int(grid_size_z), int(grid_size_r)
    extent = f"0 {nz - 1} 0 {nr - 1} 0 0"
    names = " ".join(f'ComponentName{i}="{name}"' for i, name in enumerate(field_names_list))
    xml = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" '
        'header_type="UInt32">\n'
        f'  <ImageData WholeExtent="{extent}" Origin="0 0 0" Spacing="1 1 1">\n'
        f'    <Piece Extent="{extent}">\n'
        '      <PointData>\n'
        f'        <DataArray type="Float64" Name="softy" NumberOfComponents="{len(cols)}" '
        f'{names} format="binary">\n'
        f'          {payload}\n'
        '        </DataArray>\n'
        '      </PointData>\n'
        '      <CellData>\n'
        '      </CellData>\n'
        '    </Piece>\n'
        '  </ImageData>\n'
        '</VTKFile>\n'
    )
    writer.SetFileName(filename)
    with open(filename, "w") as fh:
        fh.write(xml)


def read_vti(filename):
    """Reads a file written by vtk_write back: ({component name: (nr, nz)
    array}, (grid_size_z, grid_size_r))."""
    import xml.etree.ElementTree as ET

    root = ET.parse(filename).getroot()
    img = root.find("ImageData")
    ext = [int(v) for v in img.get("WholeExtent").split()]
    nz, nr = ext[1] + 1, ext[3] + 1
    arr = img.find("Piece/PointData/DataArray")
    k = int(arr.get("NumberOfComponents"))
    blob = base64.b64decode(arr.text.strip())
    nbytes = struct.unpack("<I", blob[:4])[0]
    data = np.frombuffer(blob[4:4 + nbytes], dtype="<f8").reshape(nz * nr, k)
    names = [arr.get(f"ComponentName{i}") for i in range(k)]
    return {nm: data[:, i].reshape(nr, nz).copy() for i, nm in enumerate(names)}, (nz, nr)
