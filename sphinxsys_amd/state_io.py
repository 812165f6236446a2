"""Body-state output and restart snapshots of device-resident particle state.

BodyStatesRecordingToVtp  the plain-text VTK PolyData writer of the reference
    (src/shared/io_system/io_vtk.cpp:165-246 and io_vtk.hpp:11-114):
    `<Body>_<sequence>.vtp` with the physical time as field data, positions
    (Float32 x 3), then `SortedParticle_ID` and the variables to write grouped by
    type in the registry's type order (UnsignedInt, int, Real, Vecd, Matd as 9
    column-major components), vertex cells with connectivity and offsets.
    Every value after the time is printed fixed with 9 decimals (the stream
    keeps the format set for the time).  `sequence` is
    padValueWithZeros(int(t * 1e6)) or "ite_" + padValueWithZeros(iteration)
    (io_base.cpp:14-54, io_base.h:72-77).
RestartIOCK  the restart snapshot (io_base.cpp:56-135, io_base_ck.hpp:117-172,
    base_particles.cpp:167-193, base_particles.hpp:180-215): one
    `Restart_<padded step>.xml` holding `restart_time` and, per body, one
    `particle` element per particle whose attributes are the evolving
    variables, printed by the vendored tinyxml2 printer (tools/tinyxml2.cpp:
    2751-2950: four spaces per depth, `/>` closing an empty element) with the
    values of XmlParser's DataToString (tools/xml_parser.h:54-80: precision
    digits10 = 15, vectors and matrices as ", "-separated coefficients, row by
    row).  Reading sets every evolving variable of the body back and returns
    the restart time.

Both read the state from the device through the C ABI (`ck.Body.read`).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

# evolving variables of a weakly-compressible fluid body under the CK step, by
# registry type: OriginalID (base_particles.cpp:33-37), PreviousSurfaceIndicator
# (surface_indication_ck.hpp:43), VolumetricMeasure, Mass, Compression,
# CompressionRate, VolumetricMeasureRef (acoustic_step_1st_half.hpp:30-34,
# density_regularization.hpp:24), Position, Velocity, ForcePrior (force_prior_ck.cpp:13)
FLUID_EVOLVING = [
    ("OriginalID", "uint"), ("PreviousSurfaceIndicator", "int"),
    ("VolumetricMeasure", "real"), ("Mass", "real"), ("Compression", "real"), ("CompressionRate", "real"),
    ("VolumetricMeasureRef", "real"),
    ("Position", "vec"), ("Velocity", "vec"), ("ForcePrior", "vec"),
]
# a static solid wall: no evolving variable changes; its geometry is kept
SOLID_EVOLVING = [("OriginalID", "uint"), ("VolumetricMeasure", "real"), ("Position", "vec"),
                  ("NormalDirection", "vec")]

_TYPE_ORDER = {"uint": 0, "int": 1, "real": 2, "vec": 3, "mat": 4}


def pad_value_with_zeros(value, width=10):
    """io_base.h:72-77: setw(width) with fill '0'."""
    return str(value).rjust(width, "0")


def time_sequence(physical_time):
    """BaseIO::convertPhysicalTimeToString (io_base.cpp:14-18)."""
    return pad_value_with_zeros(int(physical_time * 1.0e6))


def data_to_string(value):
    """XmlParser DataToString: ostream precision 15 (default float format), or
    the coefficients of a vector / matrix row by row separated by ", "."""
    a = np.asarray(value)
    if a.ndim == 0:
        if np.issubdtype(a.dtype, np.integer):
            return str(int(a))
        return "%.15g" % float(a)
    return ", ".join("%.15g" % float(x) for x in a.reshape(-1))


def _fixed9(x):
    return "%.9f" % float(x)


class BodyStatesRecordingToVtp:
    """Writes `<Body>_<sequence>.vtp` for every registered body."""

    def __init__(self, output_folder, bodies):
        self.folder = output_folder
        self.bodies = list(bodies)          # (name, ck.Body)
        self.to_write = {name: [] for name, _ in self.bodies}

    def add_to_write(self, body_name, variable, kind):
        """kind: "uint", "int", "real", "vec" or "mat" (addToWrite<DataType>)."""
        self.to_write[body_name].append((variable, kind))

    def write_to_file(self, physical_time, iteration=None):
        seq = time_sequence(physical_time) if iteration is None else "ite_" + pad_value_with_zeros(iteration)
        paths = []
        for name, body in self.bodies:
            path = os.path.join(self.folder, f"{name}_{seq}.vtp")
            self._write_body(path, name, body, physical_time)
            paths.append(path)
        return paths

    def _write_body(self, path, name, body, physical_time):
        n = body.n
        pos = body.read("Position")
        pos3 = np.zeros((n, 3))
        pos3[:, :body.dim] = pos
        out = ['<?xml version="1.0"?>\n',
               '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian">\n', " <PolyData>\n",
               "<FieldData>\n",
               '<DataArray type="Float64"  Name="TimeValue" NumberOfTuples="1" format="ascii">\n',
               _fixed9(physical_time) + "\n", " </DataArray>\n", "</FieldData>\n",
               f'  <Piece Name ="{name}" NumberOfPoints="{n}" NumberOfVerts="{n}">\n',
               "   <Points>\n",
               '    <DataArray Name="Position" type="Float32"  NumberOfComponents="3" format="ascii">\n',
               "    " + "".join(f"{_fixed9(p[0])} {_fixed9(p[1])} {_fixed9(p[2])} " for p in pos3) + "\n",
               "    </DataArray>\n", "   </Points>\n", '   <PointData  Vectors="vector">\n',
               '    <DataArray Name="SortedParticle_ID" type="Int32" format="ascii">\n',
               "    " + "".join(f"{i} " for i in range(n)) + "\n", "    </DataArray>\n"]
        for variable, kind in sorted(self.to_write[name], key=lambda vk: _TYPE_ORDER[vk[1]]):
            data = body.read(variable)
            if kind in ("uint", "int"):
                out += [f'    <DataArray Name="{variable}" type="Int32" format="ascii">\n',
                        "    " + "".join(f"{int(v)} " for v in data) + "\n"]
            elif kind == "real":
                out += [f'    <DataArray Name="{variable}" type="Float32" format="ascii">\n',
                        "    " + "".join(f"{_fixed9(v)} " for v in data) + "\n"]
            elif kind == "vec":
                v3 = np.zeros((n, 3))
                v3[:, :body.dim] = data
                out += [f'    <DataArray Name="{variable}" type="Float32"  NumberOfComponents="3" format="ascii">\n',
                        "    " + "".join(f"{_fixed9(v[0])} {_fixed9(v[1])} {_fixed9(v[2])} " for v in v3) + "\n"]
            else:
                m3 = np.zeros((n, 3, 3))
                m3[:, :body.dim, :body.dim] = data
                cols = m3.transpose(0, 2, 1).reshape(n, 9)  # column by column
                out += [f'    <DataArray Name="{variable}" type="Float32"  NumberOfComponents="9" format="ascii">\n',
                        "    " + "".join("".join(f"{_fixed9(x)} " for x in c) for c in cols) + "\n"]
            out.append("    </DataArray>\n")
        out += ["   </PointData>\n", "   <Verts>\n",
                '    <DataArray type="Int32"  Name="connectivity"  format="ascii">\n',
                "    " + "".join(f"{i} " for i in range(n)) + "\n", "    </DataArray>\n",
                '    <DataArray type="Int32"  Name="offsets"  format="ascii">\n',
                "    " + "".join(f"{i + 1} " for i in range(n)) + "\n", "    </DataArray>\n",
                "   </Verts>\n", "  </Piece>\n", " </PolyData>\n", "</VTKFile>\n"]
        with open(path, "w") as f:
            f.write("".join(out))


def read_vtp(path):
    """Arrays of a written .vtp by name (positions under "Position"), as float64."""
    root = ET.parse(path).getroot()
    arrays = {}
    for da in root.iter("DataArray"):
        name = da.attrib.get("Name")
        comps = int(da.attrib.get("NumberOfComponents", "1"))
        vals = np.array([float(x) for x in (da.text or "").split()])
        arrays[name] = vals.reshape(-1, comps) if comps > 1 else vals
    return arrays


class RestartIOCK:
    """Restart snapshots of the evolving variables of the given bodies."""

    def __init__(self, restart_folder, bodies, evolving=None):
        self.folder = restart_folder
        self.bodies = list(bodies)          # (name, ck.Body)
        self.evolving = evolving or {}
        os.makedirs(restart_folder, exist_ok=True)

    def _variables(self, name, body):
        if name in self.evolving:
            return self.evolving[name]
        return FLUID_EVOLVING if body.kind == 0 else SOLID_EVOLVING

    def path(self, step):
        return os.path.join(self.folder, "Restart_" + pad_value_with_zeros(step) + ".xml")

    def write_to_file(self, step, physical_time):
        lines = [f'<restart_data restart_time="{data_to_string(float(physical_time))}">']
        for name, body in self.bodies:
            variables = sorted(self._variables(name, body), key=lambda vk: _TYPE_ORDER[vk[1]])
            data = [(v, body.read(v)) for v, _ in variables]
            if body.n == 0:
                lines.append(f'    <body name="{name}"/>')
                continue
            lines.append(f'    <body name="{name}">')
            for i in range(body.n):
                attrs = "".join(f' {v}="{data_to_string(d[i])}"' for v, d in data)
                lines.append(f"        <particle{attrs}/>")
            lines.append("    </body>")
        lines.append("</restart_data>")
        with open(self.path(step), "w") as f:
            f.write("\n".join(lines) + "\n")
        return self.path(step)

    def read_restart_time(self, step):
        return float(ET.parse(self.path(step)).getroot().attrib["restart_time"])

    def read_from_file(self, step):
        """Writes every evolving variable found for each body back to the device
        (the particle count must match) and returns the restart time."""
        root = ET.parse(self.path(step)).getroot()
        for name, body in self.bodies:
            el = next((b for b in root.findall("body") if b.attrib.get("name") == name), None)
            if el is None:
                raise KeyError(f"restart file has no body {name}")
            particles = el.findall("particle")
            if len(particles) != body.n:
                raise ValueError(f"{name}: {len(particles)} particles in the restart file, {body.n} in the body")
            for variable, kind in self._variables(name, body):
                raw = [p.attrib[variable] for p in particles]
                if kind in ("uint", "int"):
                    body.write(variable, np.array([int(x) for x in raw], dtype=np.int64))
                else:
                    vals = np.array([[float(c) for c in x.split(",")] for x in raw])
                    body.write(variable, vals.reshape(body.n, -1) if kind != "real" else vals.reshape(body.n))
            # SortedID follows OriginalID (particle_sort_ck.hpp: sorted_id[original_id[i]] = i)
            oid = body.read("OriginalID")
            sorted_id = np.empty(body.n, dtype=np.uint32)
            sorted_id[oid] = np.arange(body.n, dtype=np.uint32)
            body.write("SortedID", sorted_id)
        return float(root.attrib["restart_time"])
