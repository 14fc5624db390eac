"""Host-side unsteady data table format (no GPU): header of the reference's column
names, values in scientific notation with 16 digits after the point, and a
lossless round trip through the reader."""
import philip_amd as P


def test_table_round_trip(tmp_path):
    rows = [{"time": 0.1 * k, "kinetic_energy": 0.125 - 1e-7 * k * k, "enstrophy": 0.375 + 1e-9 * k,
             "incompressible_kinetic_energy": 0.1249 + 1e-12 * k} for k in range(1, 6)]
    path = tmp_path / "unsteady_data_table.txt"
    P.write_unsteady_table(path, rows)
    lines = path.read_text().splitlines()
    assert lines[0].split() == ["time", "kinetic_energy", "enstrophy", "incompressible_kinetic_energy"]
    assert lines[1].split()[1] == f"{rows[0]['kinetic_energy']:.16e}"
    back = P.read_unsteady_table(path)
    assert back == rows
