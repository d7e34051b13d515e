"""Writers of Sentinel-1 product files for the ingest tests: annotation XML
and precise-orbit EOF documents in the layout of the public product
specification (the elements S1_Initialize / ComputeDerampPara /
S1_OrbitInitialize read, Func.cpp:135-819), and an independent TIFF encoder
and IFD parser (classic little / big endian, BigTIFF; written from the TIFF
6.0 and BigTIFF layouts with struct, sharing no code with the library's
reader).
"""
import struct

import numpy as np

SOL = 299792458.0
DATE = "2026-01-03"


def utc(seconds_of_day, date=DATE):
    """Seconds of the day -> 'YYYY-MM-DDTHH:MM:SS.ffffff'."""
    us = int(round(seconds_of_day * 1e6))
    h, rem = divmod(us, 3600 * 10**6)
    m, rem = divmod(rem, 60 * 10**6)
    s, f = divmod(rem, 10**6)
    return f"{date}T{h:02d}:{m:02d}:{s:02d}.{f:06d}"


def fr(x):
    """Shortest repr of x as a Python float."""
    return repr(float(x))


def seconds(u):
    """utc2seconds (Func.cpp:629-635) in the same operation order."""
    hms = u.split("T")[1]
    hh, mm, ss = hms.split(":")
    return float(hh) * 60 * 60 + float(mm) * 60 + float(ss)


def annotation_xml(sw, *, pol="VV", valid_first=None, valid_last=None, fvs_value=10,
                   lvs_value=None, geo_per_line=6, dc=None, dc_method="Data Analysis", fm=None,
                   fm_form="polynomial", decorate=False, grid_ll=None):
    """Annotation XML text of swath `sw` (a synth.Swath).

    valid_first / valid_last: per burst, first and last line with valid
    samples (firstValidSample / lastValidSample = -1 outside); dc: list of
    (az_time, t0, c0, c1, c2) estimates; fm: same for FM rates; grid_ll:
    (azimuth time, slant range m) -> (lat, lon) of the geolocation grid
    points (a fixed made-up grid when None)."""
    nb, lpb, spb = sw.n_bursts, sw.lines_per_burst, sw.samples_per_burst
    if valid_first is None:
        valid_first = list(sw.first_valid_line)
    if valid_last is None:
        valid_last = list(sw.last_valid_line)
    if lvs_value is None:
        lvs_value = spb - 11
    radar_frequency = SOL / sw.wavelength
    dt = sw.az_time_interval
    e = []
    a = e.append
    a('<?xml version="1.0" encoding="UTF-8"?>')
    if decorate:
        a("<!-- written for the ingest tests -->")
    a("<product>")
    a(f"  <adsHeader><missionId>S1A</missionId><polarisation>{pol}</polarisation></adsHeader>")
    a("  <generalAnnotation>")
    a("    <productInformation>")
    a(f"      <radarFrequency>{fr(radar_frequency)}</radarFrequency>")
    a(f"      <rangeSamplingRate>{fr(sw.range_sampling_rate)}</rangeSamplingRate>")
    a(f"      <azimuthSteeringRate>{fr(sw.steering_rate_deg)}</azimuthSteeringRate>")
    a("      <platformHeading>-1.2345678e+01</platformHeading>")
    a("    </productInformation>")
    fm = fm if fm is not None else [(sw.burst_first_line_time[b],) + tuple(sw.fm[b])
                                    for b in range(nb)]
    a(f'    <azimuthFmRateList count="{len(fm)}">')
    for (t, t0, c0, c1, c2) in fm:
        a("      <azimuthFmRate>")
        a(f"        <azimuthTime>{utc(t)}</azimuthTime><t0>{fr(t0)}</t0>")
        if fm_form == "polynomial":
            a(f'        <azimuthFmRatePolynomial count="3">{fr(c0)} {fr(c1)} {fr(c2)}'
              "</azimuthFmRatePolynomial>")
        else:
            a(f"        <c0>{fr(c0)}</c0><c1>{fr(c1)}</c1><c2>{fr(c2)}</c2>")
        a("      </azimuthFmRate>")
    a("    </azimuthFmRateList>")
    a("  </generalAnnotation>")
    a("  <imageAnnotation>")
    a("    <imageInformation>")
    a(f"      <productFirstLineUtcTime>{utc(sw.first_line_time)}</productFirstLineUtcTime>")
    a(f"      <productLastLineUtcTime>{utc(sw.last_line_time)}</productLastLineUtcTime>")
    a(f"      <ascendingNodeTime>{utc(sw.first_line_time - 1234.5)}</ascendingNodeTime>")
    a(f"      <slantRangeTime>{fr(2 * sw.slr_time_first_pixel)}</slantRangeTime>")
    a(f"      <rangePixelSpacing>{fr(sw.range_spacing)}</rangePixelSpacing>")
    a("      <azimuthPixelSpacing>1.396963e+01</azimuthPixelSpacing>")
    a(f"      <azimuthTimeInterval>{fr(dt)}</azimuthTimeInterval>")
    a(f"      <numberOfSamples>{spb}</numberOfSamples>")
    a(f"      <numberOfLines>{nb * lpb}</numberOfLines>")
    a("    </imageInformation>")
    a("    <processingInformation>")
    a(f"      <dcMethod>{dc_method}</dcMethod>")
    a("      <swathProcParamsList count=\"1\"><swathProcParams>")
    a(f"        <azimuthProcessing><processingBandwidth>{fr(sw.azimuth_bandwidth)}"
      "</processingBandwidth></azimuthProcessing>")
    a("        <rangeProcessing><processingBandwidth>5.616665e+07</processingBandwidth>"
      "</rangeProcessing>")
    a("      </swathProcParams></swathProcParamsList>")
    a("    </processingInformation>")
    a("  </imageAnnotation>")
    dc = dc if dc is not None else [
        (0.5 * (sw.burst_first_line_time[b] + sw.burst_last_line_time[b]),) + tuple(sw.dc[b])
        for b in range(nb)]
    a("  <dopplerCentroid>")
    a(f'    <dcEstimateList count="{len(dc)}">')
    for (t, t0, c0, c1, c2) in dc:
        a("      <dcEstimate>")
        a(f"        <azimuthTime>{utc(t)}</azimuthTime><t0>{fr(t0)}</t0>")
        a(f"        <geometryDcPolynomial count=\"3\">{fr(c0 + 1.0)} {fr(c1)} {fr(c2)}"
          "</geometryDcPolynomial>")
        a(f'        <dataDcPolynomial count="3">{fr(c0)} {fr(c1)} {fr(c2)}</dataDcPolynomial>')
        a("      </dcEstimate>")
    a("    </dcEstimateList>")
    a("  </dopplerCentroid>")
    a("  <swathTiming>")
    a(f"    <linesPerBurst>{lpb}</linesPerBurst>")
    a(f"    <samplesPerBurst>{spb}</samplesPerBurst>")
    a(f'    <burstList count="{nb}">')
    for b in range(nb):
        f = ["-1"] * lpb
        l_ = ["-1"] * lpb
        for y in range(lpb):
            if valid_first[b] <= y <= valid_last[b]:
                f[y] = str(fvs_value + (y % 3))
                l_[y] = str(lvs_value - (y % 2))
        a("      <burst>")
        if decorate:
            a(f"        <!-- burst {b} -->")
        a(f"        <azimuthTime>{utc(sw.burst_first_line_time[b])}</azimuthTime>")
        a(f'        <firstValidSample count="{lpb}">{" ".join(f)}</firstValidSample>')
        a(f'        <lastValidSample count="{lpb}">{" ".join(l_)}</lastValidSample>')
        a("      </burst>")
    a("    </burstList>")
    a("  </swathTiming>")
    # geolocation grid: one grid line at each burst start and one after the
    # last burst, geo_per_line points across the swath
    npts = (nb + 1) * geo_per_line
    a("  <geolocationGrid>")
    a(f"    <geolocationGridPointList count='{npts}'>")
    for g in range(nb + 1):
        line = g * lpb
        t = (sw.burst_first_line_time[g] if g < nb else sw.burst_last_line_time[nb - 1] + dt)
        for k in range(geo_per_line):
            px = round(k * (spb - 1) / (geo_per_line - 1))
            if grid_ll is None:
                lat = 30.0 + 0.1 * g - 0.01 * k
                lon = 100.0 + 0.02 * k + 0.003 * g
            else:
                lat, lon = grid_ll(t, (sw.slr_time_first_pixel + px * sw.range_spacing / SOL) * SOL)
            inc = 30.0 + 0.5 * k
            a("      <geolocationGridPoint>")
            a(f"        <azimuthTime>{utc(t)}</azimuthTime>"
              f"<slantRangeTime>{fr(2 * (sw.slr_time_first_pixel + px * sw.range_spacing / SOL))}"
              "</slantRangeTime>")
            a(f"        <line>{line}</line><pixel>{px}</pixel>"
              f"<latitude>{fr(lat)}</latitude><longitude>{fr(lon)}</longitude>"
              f"<height>0</height><incidenceAngle>{fr(inc)}</incidenceAngle>")
            a("      </geolocationGridPoint>")
    a("    </geolocationGridPointList>")
    a("  </geolocationGrid>")
    if decorate:
        a("  <note><![CDATA[<not> parsed & kept as text]]></note>")
    a("</product>")
    return "\n".join(e) + "\n"


def eof_xml(t_first, t_last, positions, velocities, date=DATE):
    """EOF document with state vectors every 10 s from t_first to t_last
    (seconds of the day); positions(t), velocities(t) -> 3-tuples."""
    e = ['<?xml version="1.0" ?>', "<Earth_Explorer_File>", "  <Earth_Explorer_Header/>",
         "  <Data_Block type=\"xml\">"]
    ts = np.arange(t_first, t_last + 1e-9, 10.0)
    e.append(f'    <List_of_OSVs count="{len(ts)}">')
    for t in ts:
        x, y, z = positions(t)
        vx, vy, vz = velocities(t)
        e.append("      <OSV>")
        e.append(f"        <TAI>TAI={utc(t + 37, date)}</TAI>")
        e.append(f"        <UTC>UTC={utc(t, date)}</UTC>")
        e.append("        <Absolute_Orbit>+12345</Absolute_Orbit>")
        e.append(f'        <X unit="m">{x:.6f}</X><Y unit="m">{y:.6f}</Y>'
                 f'<Z unit="m">{z:.6f}</Z>')
        e.append(f'        <VX unit="m/s">{vx:.6f}</VX><VY unit="m/s">{vy:.6f}</VY>'
                 f'<VZ unit="m/s">{vz:.6f}</VZ>')
        e.append("        <Quality>NOMINAL</Quality>")
        e.append("      </OSV>")
    e += ["    </List_of_OSVs>", "  </Data_Block>", "</Earth_Explorer_File>"]
    return "\n".join(e) + "\n", ts


# ---- TIFF -------------------------------------------------------------------
# (SampleFormat, BitsPerSample) of each GDAL type code
TIFF_TYPES = {8: (5, 32, np.dtype([("re", "<i2"), ("im", "<i2")])),
              6: (3, 32, np.dtype("<f4")), 7: (3, 64, np.dtype("<f8")),
              10: (6, 64, np.dtype("<c8"))}


def write_tiff(path, raster, gdt, *, big_endian=False, bigtiff=False, rows_per_strip=1,
               strip_gap=0):
    """Stripped uncompressed TIFF of `raster` ([lines, pixels] in the numpy
    layout of TIFF_TYPES[gdt], little endian in memory); strip_gap bytes of
    padding between strips."""
    fmt_code, bits, _ = TIFF_TYPES[gdt]
    lines, pixels = raster.shape[:2]
    e = ">" if big_endian else "<"
    comp = {8: 2, 6: 4, 7: 8, 10: 4}[gdt]
    body = np.ascontiguousarray(raster).view(np.uint8).reshape(lines, -1)
    if big_endian:
        body = body.reshape(lines, -1, comp)[:, :, ::-1].reshape(lines, -1)
    nstrips = (lines + rows_per_strip - 1) // rows_per_strip
    header = 16 if bigtiff else 8
    data = bytearray()
    offsets = []
    pos = header
    for s in range(nstrips):
        chunk = body[s * rows_per_strip:(s + 1) * rows_per_strip].tobytes()
        offsets.append(pos)
        data += chunk + b"\0" * strip_gap
        pos += len(chunk) + strip_gap
    counts = [len(body[s * rows_per_strip:(s + 1) * rows_per_strip].tobytes())
              for s in range(nstrips)]
    ifd_off = pos + (pos % 2)
    if ifd_off != pos:
        data += b"\0"
    ent_fmt = (e + "HHQQ") if bigtiff else (e + "HHII")
    lfmt, ltype = ("Q", 16) if bigtiff else ("I", 4)
    field = 8 if bigtiff else 4
    tags = [(256, 4, [pixels]), (257, 4, [lines]), (258, 3, [bits]), (259, 3, [1]),
            (262, 3, [1]), (273, ltype, offsets), (277, 3, [1]),
            (278, 4, [rows_per_strip]), (279, ltype, counts), (284, 3, [1]),
            (339, 3, [fmt_code])]
    n = len(tags)
    ifd_size = (8 if bigtiff else 2) + n * (20 if bigtiff else 12) + (8 if bigtiff else 4)
    extra_off = ifd_off + ifd_size
    extra = bytearray()
    ents = bytearray()
    for tag, typ, vals in tags:
        tf = {3: "H", 4: "I", 16: "Q"}[typ]
        packed = struct.pack(e + tf * len(vals), *vals)
        if len(packed) <= field:
            val = packed + b"\0" * (field - len(packed))
            ents += struct.pack(e + "HH" + ("Q" if bigtiff else "I"), tag, typ, len(vals)) + val
        else:
            ents += struct.pack(ent_fmt, tag, typ, len(vals), extra_off + len(extra))
            extra += packed
            if len(extra) % 2:
                extra += b"\0"
    ifd = struct.pack(e + ("Q" if bigtiff else "H"), n) + ents + struct.pack(e + lfmt, 0)
    if bigtiff:
        hdr = (b"MM" if big_endian else b"II") + struct.pack(e + "HHHQ", 43, 8, 0, ifd_off)
    else:
        hdr = (b"MM" if big_endian else b"II") + struct.pack(e + "HI", 42, ifd_off)
    with open(path, "wb") as f:
        f.write(hdr + bytes(data) + ifd + bytes(extra))


def parse_tiff(path):
    """Tags of the first IFD (tag -> list of values) and the raw file."""
    raw = open(path, "rb").read()
    e = "<" if raw[:2] == b"II" else ">"
    magic = struct.unpack(e + "H", raw[2:4])[0]
    big = magic == 43
    off = struct.unpack(e + "Q", raw[8:16])[0] if big else struct.unpack(e + "I", raw[4:8])[0]
    n = struct.unpack(e + ("Q" if big else "H"), raw[off:off + (8 if big else 2)])[0]
    p = off + (8 if big else 2)
    tags = {}
    for _ in range(n):
        if big:
            tag, typ, cnt = struct.unpack(e + "HHQ", raw[p:p + 12])
            field, fsz = raw[p + 12:p + 20], 8
        else:
            tag, typ, cnt = struct.unpack(e + "HHI", raw[p:p + 8])
            field, fsz = raw[p + 8:p + 12], 4
        tf, tsz = {3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}[typ]
        if cnt * tsz <= fsz:
            vals = list(struct.unpack(e + tf * cnt, field[:cnt * tsz]))
        else:
            vo = struct.unpack(e + ("Q" if big else "I"), field)[0]
            vals = list(struct.unpack(e + tf * cnt, raw[vo:vo + cnt * tsz]))
        tags[tag] = vals
        p += 20 if big else 12
    return tags, raw
