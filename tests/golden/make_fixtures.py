"""Copy the reference's data files used as fixtures (data only, no source).

  fasta/*.fasta.txt                          <- Protein_FASTA/*.fasta.txt
  Normalized_value_quality_computation.txt   <- src/Normalized_value_quality_computation.txt
                                                (200 normalised objective rows of
                                                 protein B4TWR7 printed by main.cu:839-844)
Run in the container that has the reference checkout.
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    os.makedirs(os.path.join(HERE, "fasta"), exist_ok=True)
    src = os.path.join(REF, "Protein_FASTA")
    for f in sorted(os.listdir(src)):
        shutil.copyfile(os.path.join(src, f), os.path.join(HERE, "fasta", f))
    shutil.copyfile(os.path.join(REF, "src", "Normalized_value_quality_computation.txt"),
                    os.path.join(HERE, "Normalized_value_quality_computation.txt"))


if __name__ == "__main__":
    main()
