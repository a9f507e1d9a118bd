"""Field output: VTK XML structured-grid (.vts) and raw float32 dumps.

The reference writes one field per time step through VTK
(RegularGrid.output_step / save_field_block, opesci/regulargrid.py:702-719,
and opesci_dump_field_vts_3d, src/opesciIO.cpp:614-667): point (i, j, k) of
a [dim1][dim2][dim3] array sits at ((i-m) dx, (j-m) dy, (k-m) dz) and the
values follow the array order (k fastest).  VTK is not needed here: the file
is written directly in the XML format with base64 "binary" data arrays
(UInt32 byte-count header) and no compression; ParaView and VTK readers load
it.  The reference passes {dim1, dim1, dim1} as the dimensions (correct only
for cubes); the true dimensions are written here.
"""
import base64
import struct

import numpy as np

__all__ = ['write_vts', 'read_vts', 'write_raw']


def _b64(arr):
    raw = np.ascontiguousarray(arr, dtype='<f4').tobytes()
    return base64.b64encode(struct.pack('<I', len(raw)) + raw).decode('ascii')


def write_vts(name, field, spacing, margin, field_name='field'):
    """Write `field` (dim1, dim2, dim3) to name + '.vts'; returns the path."""
    field = np.asarray(field, dtype=np.float32)
    d1, d2, d3 = field.shape
    i, j, k = np.meshgrid(np.arange(d1), np.arange(d2), np.arange(d3), indexing='ij')
    pts = np.stack([(i - margin) * np.float32(spacing[0]), (j - margin) * np.float32(spacing[1]),
                    (k - margin) * np.float32(spacing[2])], axis=-1).astype(np.float32)
    # VTK orders points with its first extent fastest: array axis 3 (k) first
    ext = '0 %d 0 %d 0 %d' % (d3 - 1, d2 - 1, d1 - 1)
    path = name + '.vts'
    with open(path, 'w') as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="StructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n')
        f.write('  <StructuredGrid WholeExtent="%s">\n' % ext)
        f.write('    <Piece Extent="%s">\n' % ext)
        f.write('      <PointData Scalars="%s">\n' % field_name)
        f.write('        <DataArray type="Float32" Name="%s" format="binary">%s</DataArray>\n'
                % (field_name, _b64(field.reshape(-1))))
        f.write('      </PointData>\n')
        f.write('      <Points>\n')
        f.write('        <DataArray type="Float32" NumberOfComponents="3" format="binary">%s</DataArray>\n'
                % _b64(pts.reshape(-1)))
        f.write('      </Points>\n')
        f.write('    </Piece>\n')
        f.write('  </StructuredGrid>\n')
        f.write('</VTKFile>\n')
    return path


def read_vts(path):
    """(field values, points) of a file written by write_vts, both float32
    in array order."""
    import re
    text = open(path).read()
    arrays = re.findall(r'format="binary">([^<]*)</DataArray>', text)
    out = []
    for a in arrays:
        raw = base64.b64decode(a)
        n = struct.unpack('<I', raw[:4])[0]
        out.append(np.frombuffer(raw[4:4 + n], dtype='<f4'))
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', text).group(1).split()]
    d3, d2, d1 = ext[1] + 1, ext[3] + 1, ext[5] + 1
    return out[0].reshape(d1, d2, d3), out[1].reshape(d1, d2, d3, 3)


def write_raw(name, field):
    """Raw float32 dump of `field` in array order to name + '.raw'."""
    path = name + '.raw'
    np.ascontiguousarray(field, dtype=np.float32).tofile(path)
    return path
