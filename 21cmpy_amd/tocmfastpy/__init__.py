"""Box file I/O with the interface of the reference's tocmfastpy.boxio."""
