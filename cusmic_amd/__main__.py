"""Command-line cleaner: ``python -m cusmic_amd SOURCE OUTPUT [options]``.

The options, their defaults and limits, the checks and the messages follow
the reference command (mod/cusmic/__main__.py:27-73, src/main.c:26-172):
SOURCE is one 2D FITS image read as float64, OUTPUT a new FITS file with the
cleaned pixels and a CRMASK extension; an existing OUTPUT is never
overwritten. Cleaning runs on the current GPU.
"""

from pathlib import Path

import click

from .fits import read_fits, write_fits

INPUT = click.Path(exists=True, dir_okay=False, path_type=Path)
OUTPUT = click.Path(dir_okay=False, path_type=Path)


@click.command()
@click.argument("source", type=INPUT)
@click.argument("output", type=OUTPUT)
@click.option("--error", type=INPUT, help="FITS image of 1-sigma errors; replaces the noise model.")
@click.option("--gain", type=click.FloatRange(min=0, min_open=True),
              help="Gain in electrons/ADU; needed without --error.")
@click.option("--readnoise", default=0.0, type=click.FloatRange(min=0), show_default=True,
              help="Read noise in electrons, with --gain.")
@click.option("--contrast", default=3.0, type=click.FloatRange(min=0), show_default=True,
              help="Minimum contrast of a detection against the fine structure.")
@click.option("--cr-threshold", default=5.0, type=click.FloatRange(min=0), show_default=True,
              help="Detection threshold in sigma.")
@click.option("--neighbor-threshold", default=3.0, type=click.FloatRange(min=0), show_default=True,
              help="Growth threshold in sigma.")
@click.option("--maxiter", default=4, type=click.IntRange(min=0), show_default=True,
              help="Maximum detection passes; 0 copies the image.")
def main(source, output, error, gain, readnoise, contrast, cr_threshold, neighbor_threshold, maxiter):
    """Remove cosmic rays from SOURCE and write the new FITS file OUTPUT."""
    if str(output).startswith("!"):
        raise click.ClickException("Choose a new output filename")
    if output.exists():
        raise click.ClickException(f"{output} already exists; choose a new output")
    if maxiter and error is None and gain is None:
        raise click.UsageError("Provide --error or --gain (optionally --readnoise)")
    try:
        data, header = read_fits(source, dtype="float64")
        if data.ndim != 2:
            raise ValueError("The FITS command expects one 2D image")
        noise = read_fits(error, dtype="float64")[0] if error else None
        if noise is not None and noise.shape != data.shape:
            raise ValueError("Error image must match the input shape")
        from . import remove_cosmics

        cleaned, mask = remove_cosmics(data, contrast, cr_threshold, neighbor_threshold, error=noise,
                                       effective_gain=gain, readnoise=readnoise, maxiter=maxiter)
        header.add_history("Cosmic rays removed with cusmic")
        write_fits(output, cleaned.cpu().numpy(), mask.cpu().numpy(), header=header)
    except (OSError, ValueError, TypeError, IndexError, KeyError) as exc:
        raise click.ClickException(str(exc)) from exc
    click.echo(f"Saved {output} ({int(mask.sum()):,} cosmic-ray pixels)")


if __name__ == "__main__":
    main()
